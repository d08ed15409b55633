// FP64 dynamics kernels for gfx950.
//
// Each kernel computes, point by point, the same expression as the reference CPU FP64 build, operand
// for operand (the results are compared bit for bit); the file and line of the reference kernel it
// matches is given at each.  Built with -ffp-contract=off: no multiply-add is fused, as the reference.
// Transcendental calls go to libm_port.hpp, which reproduces the host glibc's pow and sin.
//
// Launch shapes: point kernels run one thread per point with 64 consecutive i per wavefront (one
// 512-byte coalesced row segment per load); column kernels (vertical sums, the tridiagonal solve) run
// one thread per column and march in k, every load still i-contiguous across the wavefront.
#include "kernels_fp64.hpp"

#include <cstdlib>
#include <cstring>
#include <mutex>

#include "libm_port.hpp"

namespace MI_NS {

thread_local KernelTimer* g_timer = nullptr;

using namespace mi_dev;

namespace {

constexpr int kBx = 64;

// MI355X hands consecutive workgroups to its 8 XCDs in turn, and each XCD has its own L2.  Blocks
// that share cache lines (the halves of a 128-byte row, the halo rows of a stencil, neighbouring
// levels) should run on the same XCD, so every kernel renumbers its workgroups: the n-th workgroup of
// XCD x gets the n-th block of the x-th contiguous eighth of the grid.
constexpr unsigned kXcds = 8;
struct Blk {
  int x, y, z;
};
__device__ __forceinline__ Blk xcd_block() {
  const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const unsigned nb = gx * gy * gz;
  const unsigned p = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned xcd = p % kXcds, q = nb / kXcds, r = nb % kXcds;
  const unsigned l = xcd * q + (xcd < r ? xcd : r) + p / kXcds;  // a bijection of [0, nb)
  Blk b;
  b.x = int(l % gx);
  b.y = int((l / gx) % gy);
  b.z = int(l / (gx * gy));
  return b;
}

struct Range {
  int i0, i1, j0, j1, k0, k1;
  int kfast;  // the block order of a point kernel (below)
};

// The point kernels' blocks (64 i x 4 j at one level) in level-fastest order: the n-th block of an XCD's share is
// level n % gz of tile n / gz (tiles i-fastest), so the levels of one tile run back to back and the k +- 1 and
// k +- 2 rows of a vertical stencil are still in that XCD's L2 (a whole level-plane of the arrays a stencil
// kernel reads outgrows the 4 MB L2 at 425x300).  kfast = 0 keeps the i, j, level order of xcd_block.  (Groups of
// 4 to 32 tile rows walked level by level, which keep a horizontal stencil's halo rows hot, were measured 2-5 %
// slower at 425x300x50.)
__device__ __forceinline__ Blk point_block(int kfast) {
  Blk b = xcd_block();
  if (!kfast) return b;
  const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const unsigned l = unsigned(b.x) + gx * (unsigned(b.y) + gy * unsigned(b.z));
  b.z = int(l % gz);
  b.x = int((l / gz) % gx);
  b.y = int(l / (gz * gx));
  return b;
}
// the order point kernels launch with: level-fastest unless MI_BLOCK_ORDER=ijk (for timing the two)
int point_order() {
  static const int o = [] {
    const char* e = std::getenv("MI_BLOCK_ORDER");
    return (e && std::strcmp(e, "ijk") == 0) ? 0 : 1;
  }();
  return o;
}

inline dim3 grid_of(const Range& r, int by) {
  return dim3(unsigned((r.i1 - r.i0 + kBx - 1) / kBx), unsigned((r.j1 - r.j0 + by - 1) / by), unsigned(r.k1 - r.k0));
}

// point kernels run 64 x 4 threads per block (launch3a): the bound lets the heavier ones keep their
// registers instead of spilling
#define MI_POINT_KERNEL __global__ void __launch_bounds__(256)

// the scalar kernels: block z = level + nz * (index in the species list), one species per thread
#define MI_POINT_INDEX_SP(r)                                           \
  const Blk blk_ = point_block((r).kfast);                             \
  const int nk_ = (r).k1 - (r).k0;                                     \
  const int i = (r).i0 + blk_.x * kBx + int(threadIdx.x);              \
  const int j = (r).j0 + blk_.y * int(blockDim.y) + int(threadIdx.y);  \
  const int k = (r).k0 + blk_.z % nk_;                                 \
  const int sq_ = blk_.z / nk_;                                        \
  if (i >= (r).i1 || j >= (r).j1) return;

#define MI_POINT_INDEX(r)                          \
  const Blk blk_ = point_block((r).kfast);         \
  const int i = (r).i0 + blk_.x * kBx + int(threadIdx.x); \
  const int j = (r).j0 + blk_.y * int(blockDim.y) + int(threadIdx.y); \
  const int k = (r).k0 + blk_.z;                   \
  if (i >= (r).i1 || j >= (r).j1) return;

// HIP-event timing of the launches of one kernel (KernelTimer): events bracket each launch of the
// selected kernel on the stream it runs on, with the algorithmic bytes of that launch
bool timing(const char* name) { return g_timer && std::strcmp(name, g_timer->name) == 0; }
void time_begin(hipStream_t st) {
  KernelTimer& t = *g_timer;
  if (t.used + 2 > t.ev.size()) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    t.ev.push_back(a);
    t.ev.push_back(b);
  }
  (void)hipEventRecord(t.ev[t.used], st);
}
void time_end(hipStream_t st, double bytes) {
  KernelTimer& t = *g_timer;
  (void)hipEventRecord(t.ev[t.used + 1], st);
  t.used += 2;
  t.bytes += bytes;
}

// algorithmic bytes of one point of a kernel: distinct 3-D arrays it reads and writes, the bulk ones in the
// storage type (8 bytes in FP64, 4 in FP32-bulk) and the FP64 islands (php, the tridiagonal coefficients) in 8
// (pointwise-recomputable inputs and 1-D / 2-D arrays are not counted)
double point_bytes(const char* name) {
  struct E { const char* n; int bulk, f64; };
  static const E table[] = {
      {"large_prep", 10, 1},  // u v w al ph -> ru rv rw alt ww, php
      {"tend_u", 10, 1},      // u ru rv ww p ph alt al rw php -> ru_tend
      {"tend_v", 10, 1},      // v ru rv ww p ph alt al rw php -> rv_tend
      {"tend_w", 8, 0},       // w ru rv ww p u v -> rw_tend
      {"tend_t", 5, 0},       // t ru rv ww -> t_tend
      {"tend_ph", 6, 0},      // ph ww w u v -> ph_tend
      {"ss_prep", 27, 0},     // u v w t ph (_1 and _2) p alt ww -> u v w t ph, their 5 saves, c2a ww1, ucs vcs
      {"ss_column", 11, 3},   // ph alt t t_save c2a -> al p pm1, the damping weight, cwc cq1, a alpha gamma
      {"uv", 10, 1},          // u v ru_tend rv_tend p al alt ph php -> u v
      // u v ucs vcs t_save t_tend t ww1 ww_m ru_m rv_m ph_tend w ph_save ph rw_tend c2a alt w_save pm1 cq1 cwc
      // and the damping weight, a alpha gamma -> ww ww_m ru_m rv_m t w ph al p pm1
      {"acoustic", 33, 3},
      {"finish", 20, 0},      // u v w t ph ww and the six saves -> u v w t ph ww, al p
      {"p_rho_phi", 4, 0},    // ph t -> al p
  };
  for (const E& e : table)
    if (std::strcmp(e.n, name) == 0) return double(sizeof(real)) * e.bulk + 8.0 * e.f64;
  return 0.0;
}

template <class K, class... T>
void launch3a(const char* name, double arrays, K kern, Range r, hipStream_t st, T... args) {
  if (r.i1 <= r.i0 || r.j1 <= r.j0 || r.k1 <= r.k0) return;
  const int by = (r.j1 - r.j0) == 1 ? 1 : 4;
  r.kfast = point_order();
  const bool t = timing(name);
  if (t) time_begin(st);
  hipLaunchKernelGGL(kern, grid_of(r, by), dim3(kBx, by, 1), 0, st, r, args...);
  if (t) time_end(st, double(sizeof(real)) * arrays * double(r.i1 - r.i0) * double(r.j1 - r.j0) * double(r.k1 - r.k0));
}
// a scalar kernel over the points of r for each of nsp species (grid z = levels x species)
template <class K, class... T>
void launch3s(const char* name, double arrays, K kern, Range r, int nsp, hipStream_t st, T... args) {
  if (r.i1 <= r.i0 || r.j1 <= r.j0 || r.k1 <= r.k0 || nsp <= 0) return;
  const int by = (r.j1 - r.j0) == 1 ? 1 : 4;
  r.kfast = point_order();
  dim3 gr = grid_of(r, by);
  gr.z *= unsigned(nsp);
  const bool t = timing(name);
  if (t) time_begin(st);
  hipLaunchKernelGGL(kern, gr, dim3(kBx, by, 1), 0, st, r, args...);
  if (t) time_end(st, double(sizeof(real)) * arrays * double(r.i1 - r.i0) * double(r.j1 - r.j0) * double(r.k1 - r.k0));
}
template <class K, class... T>
void launch3(const char* name, K kern, const Range& r, hipStream_t st, T... args) {
  launch3a(name, point_bytes(name) / double(sizeof(real)), kern, r, st, args...);
}
template <class K, class... T>
void launch_cols(const char* name, K kern, int nx, int ny, int nz, hipStream_t st, T... args) {
  const bool t = timing(name);
  if (t) time_begin(st);
  hipLaunchKernelGGL(kern, dim3(unsigned((nx + kBx - 1) / kBx), unsigned(ny), 1), dim3(kBx, 1, 1), 0, st, args...);
  if (t) time_end(st, point_bytes(name) * double(nx) * double(ny) * double(nz));
}

// ----------------------------------------------------------------------------- halo fill
// One halo slot of one direction (reference semantics src/core/halo.cpp:52-84): slot a = 0..7 maps to
// i = 0, -1, -2, -3, n, n+1, n+2, n+3 on face-staggered points and -1, -2, -3, n, n+1, n+2 otherwise.
// bc = MI_BC_SPECIFIED or MI_BC_OPEN: the outermost row copied outward (halo.cpp:67-69, WRF set_physical_bc*).
template <class T>
__device__ __forceinline__ void halo_slot(T* base, idx_t sn, int a, int n, bool face, int bc, int par) {
  const int nl = face ? 4 : 3;
  const int i = a < nl ? (face ? -a : -1 - a) : n + (a - nl);
  if (flat_bc(bc)) {
    const int last = face ? n : n - 1;
    const int src = i < 0 ? 0 : (i > last ? last : i);
    if (src != i) base[idx_t(i) * sn] = base[idx_t(src) * sn];
    return;
  }
  const bool periodic = bc == MI_BC_PERIODIC;
  int src;
  if (periodic) src = ((i % n) + n) % n;
  else if (face) src = i <= 0 ? -i : 2 * n - i;
  else src = i < 0 ? -1 - i : 2 * n - 1 - i;
  T v = base[idx_t(src) * sn];
  if (!periodic && src == i) {
    if (par == 2) base[idx_t(i) * sn] = -v;
    return;
  }
  if (!periodic && par != 0) v = -v;
  base[idx_t(i) * sn] = v;
}

// x pass: thread (slot, row j, level k, field); rows of the field's interior
__global__ void k_halo_x(Geo g, HaloSet s) {
  const int a = int(threadIdx.x) & 7;
  const int j = int(blockIdx.x) * 32 + (int(threadIdx.x) >> 3);
  const int fk = int(blockIdx.y);
  const int f = fk / s.nkmax, k = fk - f * s.nkmax;
  if (f >= s.n || k >= s.nzs[f] || j >= s.nys[f]) return;
  const bool face = s.face_x[f] != 0;
  if (a >= (face ? 8 : 6)) return;
  if (s.f32[f]) halo_slot(static_cast<float*>(s.p[f]) + g.at(0, j, k), 1, a, g.nx, face, g.bcx, s.px[f]);
  else halo_slot(static_cast<double*>(s.p[f]) + g.at(0, j, k), 1, a, g.nx, face, g.bcx, s.px[f]);
}
// y pass: thread (x index over the whole row incl. halos, slot, level, field)
__global__ void k_halo_y(Geo g, HaloSet s) {
  const int i = -kHalo + int(blockIdx.x) * 64 + int(threadIdx.x);
  const int a = int(threadIdx.y);
  const int fk = int(blockIdx.y);
  const int f = fk / s.nkmax, k = fk - f * s.nkmax;
  if (f >= s.n || k >= s.nzs[f] || i >= s.nxs[f] + kHalo) return;
  const bool face = s.face_y[f] != 0;
  if (a >= (face ? 8 : 6)) return;
  if (s.f32[f]) halo_slot(static_cast<float*>(s.p[f]) + g.at(i, 0, k), g.sj, a, g.ny, face, g.bcy, s.py[f]);
  else halo_slot(static_cast<double*>(s.p[f]) + g.at(i, 0, k), g.sj, a, g.ny, face, g.bcy, s.py[f]);
}

// Every halo point of the fields of s in one pass (the reference's x pass then y pass, halo.cpp:114-131,
// composed): a halo point (hi, hj) takes the value of the point its x slot maps to in the row its y slot maps
// to, with the product of the two signs.  The x strips cover the interior rows that keep their values in y,
// the y strips every x of the row, the x halo included, as the two passes (a periodic face row n, which takes
// row 0, belongs to the y strips only, so no point is written twice with different values).  Not for a
// symmetric OddFace field, whose wall face changes sign in place during the x pass (launch_halo keeps the
// two passes for it).
// src_of: the source index of slot position i of one direction and whether the copy negates; i itself when
// the slot is a point of the domain that keeps its value
__device__ __forceinline__ int src_of(int i, int n, bool face, int bc, int par, bool& neg) {
  neg = false;
  const int last = face ? n : n - 1;
  if (flat_bc(bc)) return i < 0 ? 0 : (i > last ? last : i);
  if (bc == MI_BC_PERIODIC) return ((i % n) + n) % n;  // a face slot n takes point 0
  if (i >= 0 && i <= last) return i;
  neg = par != 0;
  if (face) return i < 0 ? -i : 2 * n - i;
  return i < 0 ? -1 - i : 2 * n - 1 - i;
}
template <class T>
__device__ __forceinline__ void halo_point(T* base, const Geo& g, int hi, int hj, bool fx, bool fy, int px, int py) {
  bool nx_, ny_;
  const int sx = src_of(hi, g.nx, fx, g.bcx, px, nx_);
  const int sy = g.two_d ? hj : src_of(hj, g.ny, fy, g.bcy, py, ny_);
  if (g.two_d) ny_ = false;
  if (sx == hi && sy == hj) return;
  const T v = base[sx + sy * g.sj];
  base[hi + hj * g.sj] = (nx_ != ny_) ? T(-v) : v;
}
__global__ void k_halo_xy(Geo g, HaloSet s) {
  const int fk = int(blockIdx.y);
  const int f = fk / s.nkmax, k = fk - f * s.nkmax;
  if (f >= s.n || k >= s.nzs[f]) return;
  const bool fx = s.face_x[f] != 0, fy = s.face_y[f] != 0;
  const int nsx = fx ? 8 : 6, nsy = g.two_d ? 0 : (fy ? 8 : 6);
  const int nys = s.nys[f], w = s.nxs[f] + 2 * kHalo;
  const int e = int(blockIdx.x) * 256 + int(threadIdx.x);
  int hi, hj;
  if (e < nys * nsx) {  // x strips: slot a of interior row j
    const int a = e % nsx;
    hj = e / nsx;
    const int nl = fx ? 4 : 3;
    hi = a < nl ? (fx ? -a : -1 - a) : g.nx + (a - nl);
    bool neg;
    if (!g.two_d && src_of(hj, g.ny, fy, g.bcy, s.py[f], neg) != hj) return;
  } else {  // y strips: slot b of column hi = -3 .. nxs + 2
    const int e2 = e - nys * nsx;
    if (e2 >= w * nsy) return;
    hi = e2 % w - kHalo;
    const int b = e2 / w, nl = fy ? 4 : 3;
    hj = b < nl ? (fy ? -b : -1 - b) : g.ny + (b - nl);
  }
  const idx_t ofs = idx_t(k) * g.sk;
  if (s.f32[f]) halo_point(static_cast<float*>(s.p[f]) + ofs, g, hi, hj, fx, fy, s.px[f], s.py[f]);
  else halo_point(static_cast<double*>(s.p[f]) + ofs, g, hi, hj, fx, fy, s.px[f], s.py[f]);
}

// ----------------------------------------------------------------------------- stores
// p[at(i, j) + kofs] = v.  The halo refreshes the reference makes after a kernel run as separate batched
// launches (launch_halo); the kind names the field's halo parity there.
enum HaloKind : int { kHMass = 0, kHU = 1, kHV = 2, kHMuu = 3, kHMuv = 4 };
template <int KIND, class T>
__device__ __forceinline__ void put(T* p, const Geo& g, int i, int j, idx_t kofs, double v) {
  p[kofs + i + j * g.sj] = T(v);
}

// ----------------------------------------------------------------------------- copies
// RK3::step's _2 -> _1 copies of whole allocations, halos included (reference rk3.cpp:94-98)
__global__ void k_copy_step(Fields f, const double* base3, const double* base2, int64_t n3, int64_t n2,
                            int64_t off3_u, int64_t off3_v, int64_t off3_w, int64_t off3_t, int64_t off3_ph,
                            int64_t off2_mu, int64_t sj2, int64_t pitch2) {
  (void)base3;
  (void)base2;
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  const int which = int(blockIdx.y);
  if (which < 5) {
    if (e >= n3) return;
    real *dst, *src;
    int64_t off;
    switch (which) {
      case 0: dst = f.u1; src = f.u2; off = off3_u; break;
      case 1: dst = f.v1; src = f.v2; off = off3_v; break;
      case 2: dst = f.w1; src = f.w2; off = off3_w; break;
      case 3: dst = f.t1; src = f.t2; off = off3_t; break;
      default: dst = f.ph1; src = f.ph2; off = off3_ph; break;
    }
    (dst - off)[e] = (src - off)[e];
  } else {  // mu, a 2-D field: the first pitch values of each row of sj (n2 = rows x pitch of them)
    if (e >= n2) return;
    const int64_t q = sj2 > 0 ? (e / pitch2) * sj2 + e % pitch2 : e;
    (f.mu1 - off2_mu)[q] = (f.mu2 - off2_mu)[q];
  }
}

__global__ void k_fill(double* p, int64_t n, double v) {
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (e < n) p[e] = v;
}
// a 2-D field's points only: the first pitch values of every row of sj (its allocation has a 3-D field's size)
__global__ void k_fill2d(double* p, int64_t sj, int pitch, int64_t rows, double v) {
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (e < int64_t(pitch) * rows) p[(e / pitch) * sj + e % pitch] = v;
}

// ----------------------------------------------------------------------------- column-group kernels
// A block holds CI consecutive columns (i) of one row j and kKT threads per column along k: the
// pointwise parts run in parallel over (column, level), the recurrences of a column (sums in k order,
// the tridiagonal sweeps) run on one thread per column out of LDS.  LDS arrays are interleaved by
// column, a[k * CI + ci], so the CI threads of one level touch consecutive banks.
// A block is CI columns x KT threads per column; a thread takes levels kt, kt + KT, ...  The shape is
// chosen per kernel and grid from kColCfgs (the domain's default, or its autotuner: every shape gives
// the same bits, the sums and sweeps of a column run on one thread in k order whatever the shape).

struct ColPos {
  int ci, kt, i, j;
};
template <int CI>
__device__ __forceinline__ ColPos col_pos() {
  const Blk b = xcd_block();
  ColPos c;
  c.ci = int(threadIdx.x) % CI;
  c.kt = int(threadIdx.x) / CI;
  c.i = b.x * CI + c.ci;
  c.j = b.y;
  return c;
}
// The k-order recurrences of a column run on one thread; their operands come from LDS in blocks of
// kChain levels loaded together, so the chain waits on LDS once per block rather than once per level.
constexpr int kChain = 8;

#define MI_COLUMN_KERNEL template <int CI, int KT> __global__ void __launch_bounds__(CI * KT)
#define LDS(arr, k, ci) (arr)[(k) * CI + (ci)]

// ----------------------------------------------------------------------------- rk_step_prep
// One kernel over the (nx+1) x nyv columns of the U/V/W points (reference LargeStep::prep,
// large_step.cpp:697-733, without its halo exchanges, none of which these kernels read):
//   KMu   the full column masses (large_step.cpp:49-66), written by the k = 0 threads and recomputed
//         by the others with the same expression (identical values)
//   KPrep3 couple_momentum, calc_alt, calc_php (large_step.cpp:68-81)
//   KWW   calc_ww_cp (large_step.cpp:83-108): divergence in parallel, the sums in k order from LDS
MI_COLUMN_KERNEL k_large_prep(Args a) {
  constexpr int kKT = KT;
  extern __shared__ double lds[];
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const ColPos cp = col_pos<CI>();
  const int i = cp.i, j = cp.j, nx = g.nx, ny = g.ny, nz = g.nz;
  const bool on = i <= nx;
  const bool mass = i < nx && j < ny;
  double* dv = lds;  // nz values per column
  if (on) {
    const idx_t x2 = g.at2(i, j);
    auto mu = [&](int ii, int jj) { return f.mu2[g.at2(ii, jj)]; };
    auto mub = [&](int ii, int jj) { return f.mub[g.at2(ii, jj)]; };
    const double mut = mass ? mu(i, j) + mub(i, j) : 0.0;
    // a specified or open edge face takes its column's mass twice (large_step.cpp:55-63)
    const bool spx = g.flx(), spy = g.fly();
    double muu = 0.0, muv = 0.0;
    if (j < ny) {
      if (spx && i == 0) muu = H * (mu(0, j) + mu(0, j) + mub(0, j) + mub(0, j));
      else if (spx && i == nx) muu = H * (mu(nx - 1, j) + mu(nx - 1, j) + mub(nx - 1, j) + mub(nx - 1, j));
      else muu = H * (mu(i, j) + mu(i - 1, j) + mub(i, j) + mub(i - 1, j));
    }
    if (i < nx && j < g.nyv) {
      if (spy && j == 0) muv = H * (mu(i, 0) + mu(i, 0) + mub(i, 0) + mub(i, 0));
      else if (spy && j == ny) muv = H * (mu(i, ny - 1) + mu(i, ny - 1) + mub(i, ny - 1) + mub(i, ny - 1));
      else muv = H * (mu(i, j) + mu(i, j - 1) + mub(i, j) + mub(i, j - 1));
    }
    if (cp.kt == 0) {
      if (mass) put<kHMass>(f.mut, g, i, j, 0, mut);
      if (j < ny) put<kHMuu>(f.muu, g, i, j, 0, muu);
      if (i < nx && j < g.nyv) put<kHMuv>(f.muv, g, i, j, 0, muv);
    }
    // calc_ww_cp's own column masses (summed in another order than calc_mu_uv)
    const double mu1 = mass ? H * (mu(i + 1, j) + mub(i + 1, j) + mu(i, j) + mub(i, j)) : 0.0;
    const double mu0 = mass ? H * (mu(i, j) + mub(i, j) + mu(i - 1, j) + mub(i - 1, j)) : 0.0;
    const double mv1 = mass ? H * (mu(i, j + 1) + mub(i, j + 1) + mu(i, j) + mub(i, j)) : 0.0;
    const double mv0 = mass ? H * (mu(i, j) + mub(i, j) + mu(i, j - 1) + mub(i, j - 1)) : 0.0;
    for (int k = cp.kt; k <= nz; k += kKT) {
      const idx_t x = g.at(i, j, k);
      if (j < ny && k < nz) put<kHU>(f.ru, g, i, j, k * g.sk, f.u2[x] * (c.c1h[k] * muu + c.c2h[k]) / f.msfuy[x2]);
      if (i < nx && j < g.nyv && k < nz) put<kHV>(f.rv, g, i, j, k * g.sk, f.v2[x] * (c.c1h[k] * muv + c.c2h[k]) * f.msfvx_inv[x2]);
      if (mass) {
        put<kHMass>(f.rw, g, i, j, k * g.sk, f.w2[x] * (c.c1f[k] * mut + c.c2f[k]) / f.msfty[x2]);
        if (k < nz) {
          put<kHMass>(f.alt, g, i, j, k * g.sk, f.al[x] + f.alb(i, j, k));
          put<kHMass>(f.php, g, i, j, k * g.sk, H * (f.phb(i, j, k) + f.phb(i, j, k + 1) + f.ph2[x] + f.ph2[x + g.sk]));
          LDS(dv, k, cp.ci) =
              f.msftx[x2] * c.dnw[k] *
              (c.rdx * ((c.c1h[k] * mu1 + c.c2h[k]) * f.u2[x + 1] / f.msfuy[x2 + 1] - (c.c1h[k] * mu0 + c.c2h[k]) * f.u2[x] / f.msfuy[x2]) +
               c.rdy * ((c.c1h[k] * mv1 + c.c2h[k]) * f.v2[x + g.sj] * f.msfvx_inv[x2 + g.sj] -
                        (c.c1h[k] * mv0 + c.c2h[k]) * f.v2[x] * f.msfvx_inv[x2]));
        }
      }
    }
  }
  __syncthreads();
  if (on && mass && cp.kt == 0) {
    double dmdt = 0;
    for (int k0 = 0; k0 < nz; k0 += kChain) {
      double v[kChain];
#pragma unroll
      for (int q = 0; q < kChain; ++q) v[q] = k0 + q < nz ? LDS(dv, k0 + q, cp.ci) : 0.0;
#pragma unroll
      for (int q = 0; q < kChain; ++q)
        if (k0 + q < nz) dmdt += v[q];
    }
    double w = 0;
    put<kHMass>(f.ww, g, i, j, 0, 0.0);
    for (int k0 = 1; k0 < nz; k0 += kChain) {
      double dvv[kChain], pv[kChain];
#pragma unroll
      for (int q = 0; q < kChain; ++q)
        if (k0 + q < nz) {
          const int kk = k0 + q;
          dvv[q] = LDS(dv, kk - 1, cp.ci);
          pv[q] = c.dnw[kk - 1] * c.c1h[kk - 1] * dmdt;
        }
#pragma unroll
      for (int q = 0; q < kChain; ++q)
        if (k0 + q < nz) {
          w = w - pv[q] - dvv[q];
          put<kHMass>(f.ww, g, i, j, (k0 + q) * g.sk, w);
        }
    }
    put<kHMass>(f.ww, g, i, j, nz * g.sk, 0.0);
  }
}

// ----------------------------------------------------------------------------- rk_tendency
// KTendU (large_step.cpp:152-264): advect_u + horizontal pressure gradient + curvature, interior faces
// MapC::vxgm (large_step.cpp:121-124): the map-factor curvature term at the mass point of index y (2-D y2)
__device__ __forceinline__ double vxgm(const Args& a, idx_t y, idx_t y2) {
  const Fields& f = a.f;
  const idx_t sj = a.g.sj;
  return H * (f.u2[y] + f.u2[y + 1]) * (f.msfvx[y2 + sj] - f.msfvx[y2]) * a.c.rdy -
         H * (f.v2[y] + f.v2[y + sj]) * (f.msfuy[y2 + 1] - f.msfuy[y2]) * a.c.rdx;
}

// kern::ord_face / ord_mass (kernels_common.hpp:45-58): the horizontal advection order next to a
// specified edge for a flux between mass points f-1 and f (ord_face) or at mass point f-1 of a field
// staggered in that direction (ord_mass); 5 stands for the domain's own order
__device__ __forceinline__ int ord_face(int f, int n, bool on) {
  if (!on) return 5;
  if (f <= 1 || f >= n - 1) return 2;
  if (f == 2 || f == n - 2) return 3;
  return 5;
}
__device__ __forceinline__ int ord_mass(int f, int n, bool on) {
  if (!on) return 5;
  if (f <= 1 || f >= n) return 2;
  if (f == 2 || f == n - 1) return 3;
  return 5;
}

// Open boundaries, advect_scalar / advect_scalar_pd (kernels_common.hpp:60-81): the outermost column (row) takes,
// instead of the flux divergence, rdx*(ub*(q_old(inner)-q_old(outer)) + q*(ru(b)-ru(a))) with the inflow mean
// normal mass flux of the cell; returned without sign
template <class T, class R>
__device__ __forceinline__ double open_x_scalar(const Geo& g, const T* q, const T* qo, const R* ru, double rdx, int i,
                                                int j, int k) {
  auto Q = [&](const T* p, int ii) { return double(p[g.at(ii, j, k)]); };
  auto RU = [&](int ii) { return double(ru[g.at(ii, j, k)]); };
  const int nx = g.nx;
  if (i == 0) {
    double ub = 0.5 * (RU(0) + RU(1));
    ub = ub < 0.0 ? ub : 0.0;
    return rdx * (ub * (Q(qo, 1) - Q(qo, 0)) + Q(q, 0) * (RU(1) - RU(0)));
  }
  double ub = 0.5 * (RU(nx - 1) + RU(nx));
  ub = ub > 0.0 ? ub : 0.0;
  return rdx * (ub * (Q(qo, nx - 1) - Q(qo, nx - 2)) + Q(q, nx - 1) * (RU(nx) - RU(nx - 1)));
}
template <class T, class R>
__device__ __forceinline__ double open_y_scalar(const Geo& g, const T* q, const T* qo, const R* rv, double rdy, int i,
                                                int j, int k) {
  auto Q = [&](const T* p, int jj) { return double(p[g.at(i, jj, k)]); };
  auto RV = [&](int jj) { return double(rv[g.at(i, jj, k)]); };
  const int ny = g.ny;
  if (j == 0) {
    double vb = 0.5 * (RV(0) + RV(1));
    vb = vb < 0.0 ? vb : 0.0;
    return rdy * (vb * (Q(qo, 1) - Q(qo, 0)) + Q(q, 0) * (RV(1) - RV(0)));
  }
  double vb = 0.5 * (RV(ny - 1) + RV(ny));
  vb = vb > 0.0 ? vb : 0.0;
  return rdy * (vb * (Q(qo, ny - 1) - Q(qo, ny - 2)) + Q(q, ny - 1) * (RV(ny) - RV(ny - 1)));
}

template <bool OPEN>
MI_POINT_KERNEL k_tend_u(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const idx_t sj = g.sj, sk = g.sk;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const auto* u = f.u2;
  const double hup = a.q.hup, vup = a.q.vup;
  const bool spx = g.flx(), spy = g.fly(), sx = g.spec();
  auto fx = [&](idx_t y, int fi) {  // the flux at mass point fi - 1, between the u points y-1 and y
    const int o = ord_mass(fi, g.nx, spx);
    if (o == 2) {  // specified: the upstream normal velocity at the boundary face (large_step.cpp:160-167)
      if (fi <= 1) {
        const double ub = (sx && u[y] < 0.0) ? double(u[y]) : double(u[y - 1]);
        return acc(Q * (f.ru[y] + f.ru[y - 1]) * (u[y] + ub));
      }
      const double ub = (sx && u[y - 1] > 0.0) ? double(u[y - 1]) : double(u[y]);
      return acc(Q * (f.ru[y] + f.ru[y - 1]) * (u[y - 1] + ub));
    }
    const acc vel = acc(H * (f.ru[y] + f.ru[y - 1]));
    if (o == 3) return acc(vel * flux3<acc>(u[y - 2], u[y - 1], u[y], u[y + 1], vel, acc(hup)));
    return acc(vel * flux5<acc>(u[y - 3], u[y - 2], u[y - 1], u[y], u[y + 1], u[y + 2], vel, acc(hup)));
  };
  auto fy = [&](idx_t y, int fj) {
    const int o = ord_face(fj, g.ny, spy);
    if (o == 2) return acc(Q * (f.rv[y] + f.rv[y - 1]) * (u[y] + u[y - sj]));
    const acc vel = acc(H * (f.rv[y] + f.rv[y - 1]));
    if (o == 3) return acc(vel * flux3<acc>(u[y - 2 * sj], u[y - sj], u[y], u[y + sj], vel, acc(hup)));
    return acc(vel * flux5<acc>(u[y - 3 * sj], u[y - 2 * sj], u[y - sj], u[y], u[y + sj], u[y + 2 * sj], vel, acc(hup)));
  };
  auto fz = [&](int kk) {
    if (kk == 0 || kk == nz) return acc(0.0);
    const idx_t y = g.at(i, j, kk);
    const acc vel = acc(H * (f.ww[y - 1] + f.ww[y]));
    if (kk == 1 || kk == nz - 1) return acc(vel * (c.fnm[kk] * u[y] + c.fnp[kk] * u[y - sk]));
    return acc(vel * flux3<acc>(u[y - 2 * sk], u[y - sk], u[y], u[y + sk], -vel, acc(vup)));
  };
  auto pu = [&](int kk) { const idx_t y = g.at(i, j, kk); return f.p[y - 1] + f.p[y]; };
  auto dpn = [&](int kk) {
    if (kk == 0) return H * (c.cf1 * pu(0) + c.cf2 * pu(1) + c.cf3 * pu(2));
    if (kk == nz) return a.q.top_lid ? H * (c.cfn * pu(nz - 1) + c.cfn1 * pu(nz - 2)) : 0.0;
    return H * (c.fnm[kk] * pu(kk) + c.fnp[kk] * pu(kk - 1));
  };
  const double msfux = f.msfux[x2], msfuy = f.msfuy[x2];
  auto RU = [&](int ii, int jj) { return double(f.ru[g.at(ii, jj, k)]); };
  auto RV = [&](int ii, int jj) { return double(f.rv[g.at(ii, jj, k)]); };
  auto UU = [&](int ii, int jj) { return double(u[g.at(ii, jj, k)]); };
  if ((OPEN && g.ox()) && (i == 0 || i == g.nx)) {
    // open x boundary face (large_step.cpp:211-228): the radiation condition with phase speed cb and, unless
    // open_ys/ye, the vertical advection; no horizontal flux divergence, pressure gradient, Coriolis, curvature
    double t;
    if (i == 0) {
      double ub = RU(0, j) - kCb * (c.c1h[k] * f.mut[g.at2(0, j)] + c.c2h[k]);
      ub = ub < 0.0 ? ub : 0.0;
      t = 0.0 - c.rdx * ub * (UU(1, j) - UU(0, j));
    } else {
      double ub = RU(i, j) + kCb * (c.c1h[k] * f.mut[g.at2(i - 1, j)] + c.c2h[k]);
      ub = ub > 0.0 ? ub : 0.0;
      t = 0.0 - c.rdx * ub * (UU(i, j) - UU(i - 1, j));
    }
    if (!(OPEN && g.oy())) t -= c.rdnw[k] * (fz(k + 1) - fz(k));
    if (a.q.use_tendf) t += f.ru_tendf[x] / msfuy;  // rk_addtend_dry
    f.ru_tend[x] = t;
    return;
  }
  double t = -(msfux * c.rdx) * (fx(x + 1, i + 1) - fx(x, i));
  if (!g.two_d) {
    if (!((OPEN && g.oy()) && (j == 0 || j == g.ny - 1))) {
      t -= (msfux * c.rdy) * (fy(x + sj, j + 1) - fy(x, j));
    } else {  // advect_u open_ys / open_ye: v-flux terms of the outermost rows (large_step.cpp:193-207)
      const int imax = g.nx - 1, ip = i < imax ? i : imax, im = i - 1 > 0 ? i - 1 : 0;
      const double mrdy = msfux * c.rdy;
      double ot;
      if (j == 0) {
        const double vw = H * (RV(ip, 0) + RV(im, 0));
        const double vb = vw < 0.0 ? vw : 0.0;
        const double dvm = RV(ip, 1) - RV(ip, 0), dvp = RV(im, 1) - RV(im, 0);
        ot = mrdy * (vb * (UU(i, 1) - UU(i, 0)) + H * UU(i, 0) * (dvm + dvp));
      } else {
        const int je = g.ny;
        const double vw = H * (RV(ip, je) + RV(im, je));
        const double vb = vw > 0.0 ? vw : 0.0;
        const double dvm = RV(ip, je) - RV(ip, je - 1), dvp = RV(im, je) - RV(im, je - 1);
        ot = mrdy * (vb * (UU(i, je - 1) - UU(i, je - 2)) + H * UU(i, je - 1) * (dvm + dvp));
      }
      t -= ot;
    }
  }
  if (!((OPEN && g.oy()) && spx && (i == 0 || i == g.nx))) t -= c.rdnw[k] * (fz(k + 1) - fz(k));
  const double rm = msfux / msfuy;
  double dpx = rm * H * c.rdx * (c.c1h[k] * f.muu[x2] + c.c2h[k]) *
               ((f.ph2[x + sk] - f.ph2[x + sk - 1] + f.ph2[x] - f.ph2[x - 1]) + (f.alt[x] + f.alt[x - 1]) * (f.p[x] - f.p[x - 1]) +
                (f.al[x] + f.al[x - 1]) * (f.pb(i, j, k) - f.pb(i - 1, j, k)));
  dpx += rm * c.rdx * (f.php[x] - f.php[x - 1]) *
         (c.rdnw[k] * (dpn(k + 1) - dpn(k)) - H * (c.c1h[k] * f.mu2[x2 - 1] + c.c1h[k] * f.mu2[x2]));
  t -= (a.q.moist ? f.cqu[x] : 1.0) * dpx;  // kern::MoistCoef: 1 when dry
  // Coriolis, curvature (large_step.cpp:248-259)
  const double rv4 = Q * (f.rv[x - 1 + sj] + f.rv[x + sj] + f.rv[x - 1] + f.rv[x]);
  const double rw4 = Q * (f.rw[x - 1 + sk] + f.rw[x - 1] + f.rw[x + sk] + f.rw[x]);
  if (a.q.cor)
    t = t + rm * H * (f.fcor[x2] + f.fcor[x2 - 1]) * rv4 -
        H * (f.ecor[x2] + f.ecor[x2 - 1]) * H * (f.cosa[x2] + f.cosa[x2 - 1]) * rw4;
  if (a.q.mapped)
    t = t + H * (vxgm(a, x, x2) + vxgm(a, x - 1, x2 - 1)) * rv4 - u[x] * kReradius * rw4;
  else
    t = t - u[x] * kReradius * rw4;
  if (a.q.use_tendf) t += f.ru_tendf[x] / msfuy;  // rk_addtend_dry
  f.ru_tend[x] = t;
}

// KTendV (large_step.cpp:267-376)
template <bool OPEN>
MI_POINT_KERNEL k_tend_v(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const idx_t sj = g.sj, sk = g.sk;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const auto* v = f.v2;
  const double hup = a.q.hup, vup = a.q.vup;
  const bool spx = g.flx(), spy = g.fly(), sy = g.spec();
  auto fx = [&](idx_t y, int fi) {
    const int o = ord_face(fi, g.nx, spx);
    if (o == 2) return acc(Q * (f.ru[y] + f.ru[y - sj]) * (v[y] + v[y - 1]));
    const acc vel = acc(H * (f.ru[y] + f.ru[y - sj]));
    if (o == 3) return acc(vel * flux3<acc>(v[y - 2], v[y - 1], v[y], v[y + 1], vel, acc(hup)));
    return acc(vel * flux5<acc>(v[y - 3], v[y - 2], v[y - 1], v[y], v[y + 1], v[y + 2], vel, acc(hup)));
  };
  auto fy = [&](idx_t y, int fj) {  // the flux at mass row fj - 1
    const int o = ord_mass(fj, g.ny, spy);
    if (o == 2) {  // the upstream normal velocity at the boundary face (large_step.cpp:282-289)
      if (fj <= 1) {
        const double vb = (sy && v[y] < 0.0) ? double(v[y]) : double(v[y - sj]);
        return acc(Q * (f.rv[y] + f.rv[y - sj]) * (v[y] + vb));
      }
      const double vb = (sy && v[y - sj] > 0.0) ? double(v[y - sj]) : double(v[y]);
      return acc(Q * (f.rv[y] + f.rv[y - sj]) * (vb + v[y - sj]));
    }
    const acc vel = acc(H * (f.rv[y] + f.rv[y - sj]));
    if (o == 3) return acc(vel * flux3<acc>(v[y - 2 * sj], v[y - sj], v[y], v[y + sj], vel, acc(hup)));
    return acc(vel * flux5<acc>(v[y - 3 * sj], v[y - 2 * sj], v[y - sj], v[y], v[y + sj], v[y + 2 * sj], vel, acc(hup)));
  };
  auto fz = [&](int kk) {
    if (kk == 0 || kk == nz) return acc(0.0);
    const idx_t y = g.at(i, j, kk);
    const acc vel = acc(H * (f.ww[y - sj] + f.ww[y]));
    if (kk == 1 || kk == nz - 1) return acc(vel * (c.fnm[kk] * v[y] + c.fnp[kk] * v[y - sk]));
    return acc(vel * flux3<acc>(v[y - 2 * sk], v[y - sk], v[y], v[y + sk], -vel, acc(vup)));
  };
  auto pv = [&](int kk) { const idx_t y = g.at(i, j, kk); return f.p[y - sj] + f.p[y]; };
  auto dpn = [&](int kk) {
    if (kk == 0) return H * (c.cf1 * pv(0) + c.cf2 * pv(1) + c.cf3 * pv(2));
    if (kk == nz) return a.q.top_lid ? H * (c.cfn * pv(nz - 1) + c.cfn1 * pv(nz - 2)) : 0.0;
    return H * (c.fnm[kk] * pv(kk) + c.fnp[kk] * pv(kk - 1));
  };
  const double msfvy = f.msfvy[x2], msfvx = f.msfvx[x2];
  auto RU = [&](int ii, int jj) { return double(f.ru[g.at(ii, jj, k)]); };
  auto RV = [&](int ii, int jj) { return double(f.rv[g.at(ii, jj, k)]); };
  auto VV = [&](int ii, int jj) { return double(v[g.at(ii, jj, k)]); };
  if ((OPEN && g.oy()) && (j == 0 || j == g.ny)) {  // open y boundary face: radiation only (large_step.cpp:327-340)
    double t;
    if (j == 0) {
      double vb = RV(i, 0) - kCb * (c.c1h[k] * f.mut[g.at2(i, 0)] + c.c2h[k]);
      vb = vb < 0.0 ? vb : 0.0;
      t = 0.0 - c.rdy * vb * (VV(i, 1) - VV(i, 0));
    } else {
      double vb = RV(i, j) + kCb * (c.c1h[k] * f.mut[g.at2(i, j - 1)] + c.c2h[k]);
      vb = vb > 0.0 ? vb : 0.0;
      t = 0.0 - c.rdy * vb * (VV(i, j) - VV(i, j - 1));
    }
    if (a.q.use_tendf) t += f.rv_tendf[x] * f.msfvx_inv[x2];  // rk_addtend_dry
    f.rv_tend[x] = t;
    return;
  }
  const double rm = msfvy / msfvx;
  double t;
  if ((OPEN && g.ox()) && (i == 0 || i == g.nx - 1)) {
    // open x: the y flux, then advect_v's open_xs / open_xe term instead of the x flux (large_step.cpp:307-323)
    t = -(msfvy * c.rdy) * (fy(x + sj, j + 1) - fy(x, j));
    const int jmax = g.ny - 1, jp = j < jmax ? j : jmax, jm = j - 1 > 0 ? j - 1 : 0;
    double ot;
    if (i == 0) {
      const double mrdx = f.msfvy[g.at2(0, j)] * c.rdx;
      const double uw = H * (RU(0, jp) + RU(0, jm));
      const double ub = uw < 0.0 ? uw : 0.0;
      const double dup = RU(1, jp) - RU(0, jp), dum = RU(1, jm) - RU(0, jm);
      ot = mrdx * (ub * (VV(1, j) - VV(0, j)) + H * VV(0, j) * (dup + dum));
    } else {
      const int ie = g.nx;
      const double mrdx = f.msfvy[g.at2(ie - 1, j)] * c.rdx;
      const double uw = H * (RU(ie, jp) + RU(ie, jm));
      const double ub = uw > 0.0 ? uw : 0.0;
      const double dup = RU(ie, jp) - RU(ie - 1, jp), dum = RU(ie, jm) - RU(ie - 1, jm);
      ot = mrdx * (ub * (VV(ie - 1, j) - VV(ie - 2, j)) + H * VV(ie - 1, j) * (dup + dum));
    }
    t -= ot;
  } else {
    t = -(msfvy * c.rdx) * (fx(x + 1, i + 1) - fx(x, i));
    t -= (msfvy * c.rdy) * (fy(x + sj, j + 1) - fy(x, j));
  }
  t -= rm * c.rdnw[k] * (fz(k + 1) - fz(k));
  double dpy = rm * H * c.rdy * (c.c1h[k] * f.muv[x2] + c.c2h[k]) *
               ((f.ph2[x + sk] - f.ph2[x + sk - sj] + f.ph2[x] - f.ph2[x - sj]) + (f.alt[x] + f.alt[x - sj]) * (f.p[x] - f.p[x - sj]) +
                (f.al[x] + f.al[x - sj]) * (f.pb(i, j, k) - f.pb(i, j - 1, k)));
  dpy += rm * c.rdy * (f.php[x] - f.php[x - sj]) *
         (c.rdnw[k] * (dpn(k + 1) - dpn(k)) - H * (c.c1h[k] * f.mu2[x2 - sj] + c.c1h[k] * f.mu2[x2]));
  t -= (a.q.moist ? f.cqv[x] : 1.0) * dpy;
  // Coriolis, curvature (large_step.cpp:360-371)
  const double ru4 = Q * (f.ru[x] + f.ru[x + 1] + f.ru[x - sj] + f.ru[x + 1 - sj]);
  const double rw4 = Q * (f.rw[x - sj + sk] + f.rw[x - sj] + f.rw[x + sk] + f.rw[x]);
  if (a.q.cor)
    t = t - rm * H * (f.fcor[x2] + f.fcor[x2 - sj]) * ru4 +
        rm * H * (f.ecor[x2] + f.ecor[x2 - sj]) * H * (f.sina[x2] + f.sina[x2 - sj]) * rw4;
  if (a.q.mapped)
    t = t - H * (vxgm(a, x, x2) + vxgm(a, x - sj, x2 - sj)) * ru4 - rm * v[x] * kReradius * rw4;
  else
    t = t - rm * v[x] * kReradius * rw4;
  if (a.q.use_tendf) t += f.rv_tendf[x] * f.msfvx_inv[x2];  // rk_addtend_dry
  f.rv_tend[x] = t;
}

// KTendW (large_step.cpp:379-499), w levels 1..nz
template <bool OPEN>
MI_POINT_KERNEL k_tend_w(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const idx_t sj = g.sj, sk = g.sk;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const auto* w = f.w2;
  const double hup = a.q.hup, vup = a.q.vup;
  auto rux = [&](idx_t y) {  // ru at w level k
    if (k < nz) return c.fnm[k] * f.ru[y] + c.fnp[k] * f.ru[y - sk];
    return (2.0 - c.fnm[nz - 1]) * f.ru[y - sk] - c.fnp[nz - 1] * f.ru[y - 2 * sk];
  };
  auto rvy = [&](idx_t y) {
    if (k < nz) return c.fnm[k] * f.rv[y] + c.fnp[k] * f.rv[y - sk];
    return (2.0 - c.fnm[nz - 1]) * f.rv[y - sk] - c.fnp[nz - 1] * f.rv[y - 2 * sk];
  };
  const bool spx = g.flx(), spy = g.fly();
  auto fx = [&](idx_t y, int fi) {
    const acc vel = acc(rux(y));
    const int o = ord_face(fi, g.nx, spx);
    if (o == 2) return acc(H * vel * (w[y] + w[y - 1]));
    if (o == 3) return acc(vel * flux3<acc>(w[y - 2], w[y - 1], w[y], w[y + 1], vel, acc(hup)));
    return acc(vel * flux5<acc>(w[y - 3], w[y - 2], w[y - 1], w[y], w[y + 1], w[y + 2], vel, acc(hup)));
  };
  auto fy = [&](idx_t y, int fj) {
    const acc vel = acc(rvy(y));
    const int o = ord_face(fj, g.ny, spy);
    if (o == 2) return acc(H * vel * (w[y] + w[y - sj]));
    if (o == 3) return acc(vel * flux3<acc>(w[y - 2 * sj], w[y - sj], w[y], w[y + sj], vel, acc(hup)));
    return acc(vel * flux5<acc>(w[y - 3 * sj], w[y - 2 * sj], w[y - sj], w[y], w[y + sj], w[y + 2 * sj], vel, acc(hup)));
  };
  auto fz = [&](int m) {  // vertical flux at mass level m
    const idx_t y = g.at(i, j, m);
    if (m == 0 || m == nz - 1) return acc(0.25 * (f.ww[y + sk] + f.ww[y]) * (w[y + sk] + w[y]));
    const acc vel = acc(H * (f.ww[y + sk] + f.ww[y]));
    return acc(vel * flux3<acc>(w[y - sk], w[y], w[y + sk], w[y + 2 * sk], -vel, acc(vup)));
  };
  const double msftx = f.msftx[x2], msfty = f.msfty[x2];
  double t;
  const bool ocol = (OPEN && g.ox()) && (i == 0 || i == g.nx - 1), orow = (OPEN && g.oy()) && (j == 0 || j == g.ny - 1);
  if (ocol || orow) {
    // advect_w's open terms (large_step.cpp:416-466), WRF order: y flux, x flux, open x terms, open y terms;
    // ru (rv) at w level k from the two u (v) faces of the edge column (row), extrapolated at the lid
    auto RU = [&](int ii, int kk) { return double(f.ru[g.at(ii, j, kk)]); };
    auto RV = [&](int jj, int kk) { return double(f.rv[g.at(i, jj, kk)]); };
    auto WW = [&](int ii, int jj) { return double(w[g.at(ii, jj, k)]); };
    t = 0.0;
    if (!g.two_d && !orow) t = t - (msftx * c.rdy) * (fy(x + sj, j + 1) - fy(x, j));
    if (!ocol) t = t - (msftx * c.rdx) * (fx(x + 1, i + 1) - fx(x, i));
    if (ocol) {
      const int ea = i == 0 ? 0 : g.nx - 1, eb = i == 0 ? 1 : g.nx;
      double uw, dru;
      if (k < nz) {
        uw = H * (c.fnm[k] * (RU(ea, k) + RU(eb, k)) + c.fnp[k] * (RU(ea, k - 1) + RU(eb, k - 1)));
        dru = c.fnm[k] * (RU(eb, k) - RU(ea, k)) + c.fnp[k] * (RU(eb, k - 1) - RU(ea, k - 1));
      } else {
        const double fm = 2.0 - c.fnm[nz - 1], fp = c.fnp[nz - 1];
        uw = H * (fm * (RU(ea, nz - 1) + RU(eb, nz - 1)) - fp * (RU(ea, nz - 2) + RU(eb, nz - 2)));
        dru = fm * (RU(eb, nz - 1) - RU(ea, nz - 1)) - fp * (RU(eb, nz - 2) - RU(ea, nz - 2));
      }
      double ot;
      if (i == 0) {
        const double ub = uw < 0.0 ? uw : 0.0;
        ot = c.rdx * (ub * (WW(1, j) - WW(0, j)) + WW(0, j) * dru);
      } else {
        const double ub = uw > 0.0 ? uw : 0.0;
        ot = c.rdx * (ub * (WW(i, j) - WW(i - 1, j)) + WW(i, j) * dru);
      }
      t = t - ot;
    }
    if (orow) {
      const int ea = j == 0 ? 0 : g.ny - 1, eb = j == 0 ? 1 : g.ny;
      double vw, drv;
      if (k < nz) {
        vw = H * (c.fnm[k] * (RV(ea, k) + RV(eb, k)) + c.fnp[k] * (RV(ea, k - 1) + RV(eb, k - 1)));
        drv = c.fnm[k] * (RV(eb, k) - RV(ea, k)) + c.fnp[k] * (RV(eb, k - 1) - RV(ea, k - 1));
      } else {
        const double fm = 2.0 - c.fnm[nz - 1], fp = c.fnp[nz - 1];
        vw = H * (fm * (RV(ea, nz - 1) + RV(eb, nz - 1)) - fp * (RV(ea, nz - 2) + RV(eb, nz - 2)));
        drv = fm * (RV(eb, nz - 1) - RV(ea, nz - 1)) - fp * (RV(eb, nz - 2) - RV(ea, nz - 2));
      }
      double ot;
      if (j == 0) {
        const double vb = vw < 0.0 ? vw : 0.0;
        ot = c.rdy * (vb * (WW(i, 1) - WW(i, 0)) + WW(i, 0) * drv);
      } else {
        const double vb = vw > 0.0 ? vw : 0.0;
        ot = c.rdy * (vb * (WW(i, j) - WW(i, j - 1)) + WW(i, j) * drv);
      }
      t = t - ot;
    }
  } else {
    t = -(msftx * c.rdx) * (fx(x + 1, i + 1) - fx(x, i));
    if (!g.two_d) t -= (msftx * c.rdy) * (fy(x + sj, j + 1) - fy(x, j));
  }
  const double rmsft = 1.0 / msfty;
  if (k < nz) {
    t -= c.rdn[k] * (fz(k) - fz(k - 1));
    const double cqw = a.q.moist ? f.cqw[x] : 0.0, cq1 = 1.0 / (1.0 + cqw), cq2 = cqw * cq1;
    t += rmsft * kG * (cq1 * c.rdn[k] * (f.p[x] - f.p[x - sk]) - c.c1f[k] * f.mu2[x2] - cq2 * (c.c1f[k] * f.mub[x2] + c.c2f[k]));
    if (a.q.wd_dt != 0.0) {
      const double muf = c.c1f[k] * f.mut[x2] + c.c2f[k];
      double cfl = f.ww[x] / muf * c.rdnw[k] * a.q.wd_dt;
      cfl = cfl < 0.0 ? -cfl : cfl;
      if (cfl > 1.0) t = t - (w[x] < 0.0 ? -1.0 : 1.0) * 0.3 * (cfl - a.q.w_crit) * muf;
    }
    const double ruw = H * (c.fnm[k] * (f.ru[x] + f.ru[x + 1]) + c.fnp[k] * (f.ru[x - sk] + f.ru[x - sk + 1]));
    const double rvw = H * (c.fnm[k] * (f.rv[x] + f.rv[x + sj]) + c.fnp[k] * (f.rv[x - sk] + f.rv[x - sk + sj]));
    const double rmt = msftx / msfty;
    if (a.q.cor) t = t + f.ecor[x2] * (f.cosa[x2] * ruw - rmt * f.sina[x2] * rvw);
    const double uw = H * (c.fnm[k] * (f.u2[x] + f.u2[x + 1]) + c.fnp[k] * (f.u2[x - sk] + f.u2[x - sk + 1]));
    const double vw = H * (c.fnm[k] * (f.v2[x] + f.v2[x + sj]) + c.fnp[k] * (f.v2[x - sk] + f.v2[x - sk + sj]));
    t = t + kReradius * (ruw * uw + rmt * rvw * vw);
  } else {
    t += 2.0 * c.rdn[nz - 1] * fz(nz - 1);
    const double cqw = a.q.moist ? f.cqw[x - sk] : 0.0, cq1 = 1.0 / (1.0 + cqw), cq2 = cqw * cq1;
    t += rmsft * kG *
         (cq1 * 2.0 * c.rdnw[nz - 1] * (-f.p[x - sk]) - c.c1f[nz] * f.mu2[x2] - cq2 * (c.c1f[nz] * f.mub[x2] + c.c2f[nz]));
  }
  if (a.q.use_tendf) t += f.rw_tendf[x] / msfty;  // rk_addtend_dry
  f.rw_tend[x] = t;
}

// KTendT (large_step.cpp:502-565)
template <bool OPEN, bool THF>
MI_POINT_KERNEL k_tend_t(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const idx_t sj = g.sj, sk = g.sk;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const auto* t2 = f.t2;
  const double hup = a.q.hup, vup = a.q.vup;
  const bool spx = g.flx(), spy = g.fly();
  auto fx = [&](idx_t y, int fi) {
    const acc vel = acc(f.ru[y]);
    const int o = ord_face(fi, g.nx, spx);
    if (o == 2) return acc(H * vel * (t2[y] + t2[y - 1]));
    if (o == 3) return acc(vel * flux3<acc>(t2[y - 2], t2[y - 1], t2[y], t2[y + 1], vel, acc(hup)));
    return acc(vel * flux5<acc>(t2[y - 3], t2[y - 2], t2[y - 1], t2[y], t2[y + 1], t2[y + 2], vel, acc(hup)));
  };
  auto fy = [&](idx_t y, int fj) {
    const acc vel = acc(f.rv[y]);
    const int o = ord_face(fj, g.ny, spy);
    if (o == 2) return acc(H * vel * (t2[y] + t2[y - sj]));
    if (o == 3) return acc(vel * flux3<acc>(t2[y - 2 * sj], t2[y - sj], t2[y], t2[y + sj], vel, acc(hup)));
    return acc(vel * flux5<acc>(t2[y - 3 * sj], t2[y - 2 * sj], t2[y - sj], t2[y], t2[y + sj], t2[y + 2 * sj], vel, acc(hup)));
  };
  auto fz = [&](int kk) {
    if (kk == 0 || kk == nz) return acc(0.0);
    const idx_t y = g.at(i, j, kk);
    const acc vel = acc(f.ww[y]);
    if (kk == 1 || kk == nz - 1) return acc(vel * (c.fnm[kk] * t2[y] + c.fnp[kk] * t2[y - sk]));
    return acc(vel * flux3<acc>(t2[y - 2 * sk], t2[y - sk], t2[y], t2[y + sk], -vel, acc(vup)));
  };
  const double msftx = f.msftx[x2];
  double rr;
  const bool ocol = (OPEN && g.ox()) && (i == 0 || i == g.nx - 1), orow = (OPEN && g.oy()) && (j == 0 || j == g.ny - 1);
  if (ocol || orow) {  // open boundary: WRF order y flux, x flux, open x terms, open y terms (large_step.cpp:534-541)
    rr = 0.0;
    if (!g.two_d && !orow) rr = rr - (msftx * c.rdy) * (fy(x + sj, j + 1) - fy(x, j));
    if (!ocol) rr = rr - (msftx * c.rdx) * (fx(x + 1, i + 1) - fx(x, i));
    if (ocol) rr = rr - open_x_scalar(a.g, t2, t2, f.ru, c.rdx, i, j, k);
    if (orow) rr = rr - open_y_scalar(a.g, t2, t2, f.rv, c.rdy, i, j, k);
  } else {
    rr = -(msftx * c.rdx) * (fx(x + 1, i + 1) - fx(x, i));
    if (!g.two_d) rr -= (msftx * c.rdy) * (fy(x + sj, j + 1) - fy(x, j));
  }
  rr -= c.rdnw[k] * (fz(k + 1) - fz(k));
  if (THF && f.thf) {  // set_tend(RTHFTEN) (large_step.cpp:543-555): the advection times msfty; specified edges skip their flux
    const bool scol = g.spec() && (i == 0 || i == g.nx - 1), srow = g.spec() && (j == 0 || j == g.ny - 1);
    if (!scol && !srow) {
      f.thf[x] = rr * f.msfty[x2];
    } else {
      double q = 0.0;
      if (!scol) q = -(msftx * c.rdx) * (fx(x + 1, i + 1) - fx(x, i));
      if (!g.two_d && !srow) q -= (msftx * c.rdy) * (fy(x + sj, j + 1) - fy(x, j));
      q -= c.rdnw[k] * (fz(k + 1) - fz(k));
      f.thf[x] = q * f.msfty[x2];
    }
  }
  if (a.q.use_tendf) rr += f.t_tendf[x] / f.msfty[x2];  // rk_addtend_dry
  // rk_addtend_dry's coupled heating (large_step.cpp:557), with moisture
  if (a.q.moist) rr += (c.c1h[k] * f.mut[x2] + c.c2h[k]) * f.hdia[x] / f.msfty[x2];
  f.t_tend[x] = rr;
}

// KTendPh rhs_ph (large_step.cpp:570-668), w levels 1..nz, 6th-order horizontal gradient
template <bool OPEN>
MI_POINT_KERNEL k_tend_ph(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const idx_t sj = g.sj, sk = g.sk;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const auto* ph = f.ph2;
  const auto* u = f.u2;
  const auto* v = f.v2;
  auto wdwn = [&](int m) {
    const idx_t y = g.at(i, j, m);
    return H * (f.ww[y + sk] + f.ww[y]) * c.rdnw[m] * (ph[y + sk] - ph[y] + f.phb(i, j, m + 1) - f.phb(i, j, m));
  };
  auto grad6x = [&]() {
    return 45.0 * (ph[x + 1] - ph[x - 1]) - 9.0 * (ph[x + 2] - ph[x - 2]) + (ph[x + 3] - ph[x - 3]) +
           45.0 * (f.phb(i + 1, j, k) - f.phb(i - 1, j, k)) - 9.0 * (f.phb(i + 2, j, k) - f.phb(i - 2, j, k)) +
           (f.phb(i + 3, j, k) - f.phb(i - 3, j, k));
  };
  auto grad6y = [&]() {
    return 45.0 * (ph[x + sj] - ph[x - sj]) - 9.0 * (ph[x + 2 * sj] - ph[x - 2 * sj]) + (ph[x + 3 * sj] - ph[x - 3 * sj]) +
           45.0 * (f.phb(i, j + 1, k) - f.phb(i, j - 1, k)) - 9.0 * (f.phb(i, j + 2, k) - f.phb(i, j - 2, k)) +
           (f.phb(i, j + 3, k) - f.phb(i, j - 3, k));
  };
  auto grad4x = [&]() {
    return 8.0 * (ph[x + 1] - ph[x - 1]) - (ph[x + 2] - ph[x - 2]) + 8.0 * (f.phb(i + 1, j, k) - f.phb(i - 1, j, k)) -
           (f.phb(i + 2, j, k) - f.phb(i - 2, j, k));
  };
  auto grad4y = [&]() {
    return 8.0 * (ph[x + sj] - ph[x - sj]) - (ph[x + 2 * sj] - ph[x - 2 * sj]) +
           8.0 * (f.phb(i, j + 1, k) - f.phb(i, j - 1, k)) - (f.phb(i, j + 2, k) - f.phb(i, j - 2, k));
  };
  // one-sided differences (dfwd) of ph + phb at index y (base point (bi, bj))
  auto dfwdx = [&](idx_t y, int bi) { return f.phb(bi + 1, j, k) - f.phb(bi, j, k) + ph[y + 1] - ph[y]; };
  auto dfwdy = [&](idx_t y, int bj) { return f.phb(i, bj + 1, k) - f.phb(i, bj, k) + ph[y + sj] - ph[y]; };
  // rhs_ph next to specified or open edges (large_step.cpp:596-605): none on the outer row, 2nd order on the
  // next, then 4th order in y (in x only with open_xs/xe), 6th from the fourth row in
  const bool sp = g.spec(), ox = (OPEN && g.ox());
  auto horder = [&](int ii, int n, bool on, bool is_x) {
    if (!on) return 6;
    if (ii <= 0 || ii >= n - 1) return 0;
    if (ii == 1 || ii == n - 2) return 2;
    if (ii == 2 || ii == n - 3) return (is_x && !ox) ? 0 : 4;
    return 6;
  };
  const double msfty = f.msfty[x2];
  double t = 0;
  if (k < nz) t = t - (c.fnm[k] * wdwn(k) + c.fnp[k] * wdwn(k - 1));
  t = t + (c.c1f[k] * f.mut[x2] + c.c2f[k]) * kG * f.w2[x] / msfty;
  const double r60 = 1.0 / 60.0, r12 = 1.0 / 12.0;
  const idx_t xt = g.at(i, j, nz - 1);  // top extrapolation reads levels nz-1, nz-2
  if (!g.two_d) {
    const double mvp = c.c1f[k] * f.muv[x2 + sj] + c.c2f[k], mvm = c.c1f[k] * f.muv[x2] + c.c2f[k];
    const double vp = k < nz ? v[x + sj] + v[x + sj - sk] : c.cfn * v[xt + sj] + c.cfn1 * v[xt + sj - sk];
    const double vm = k < nz ? v[x] + v[x - sk] : c.cfn * v[xt] + c.cfn1 * v[xt - sk];
    const double cy = (k < nz ? 0.25 : H) * c.rdy / msfty;
    const int o = horder(j, g.ny, g.fly(), false);
    if (o == 6)
      t = t - cy * ((mvp * vp * f.msfvy[x2 + sj] + mvm * vm * f.msfvy[x2]) * r60 * grad6y());
    else if (o == 4)
      t = t - cy * ((mvp * vp * f.msfvy[x2 + sj] + mvm * vm * f.msfvy[x2]) * r12 * grad4y());
    else if (o == 2)
      t = t - cy * (mvp * vp * f.msfvy[x2 + sj] * dfwdy(x, j) + mvm * vm * f.msfvy[x2] * dfwdy(x - sj, j - 1));
  }
  const double mup = c.c1f[k] * f.muu[x2 + 1] + c.c2f[k], mum = c.c1f[k] * f.muu[x2] + c.c2f[k];
  const double up = k < nz ? u[x + 1] + u[x + 1 - sk] : c.cfn * u[xt + 1] + c.cfn1 * u[xt + 1 - sk];
  const double um = k < nz ? u[x] + u[x - sk] : c.cfn * u[xt] + c.cfn1 * u[xt - sk];
  const double cx = (k < nz ? 0.25 : H) * c.rdx / msfty;
  const int o = horder(i, g.nx, g.flx(), true);
  if (o == 6)
    t = t - cx * ((mup * up * f.msfux[x2 + 1] + mum * um * f.msfux[x2]) * r60 * grad6x());
  else if (o == 2)
    t = t - cx * (mup * up * f.msfux[x2 + 1] * dfwdx(x, i) + mum * um * f.msfux[x2] * dfwdx(x - 1, i - 1));
  else if (o == 4)
    t = t - cx * ((mup * up * f.msfux[x2 + 1] + mum * um * f.msfux[x2]) * r12 * grad4x());
  // open boundaries (large_step.cpp:636-661), south/north then west/east: inflow advection of ph by the normal
  // wind at the w level (the west lid level takes none: WRF's kz = kde there reads fnm/fnp = 0)
  if ((OPEN && g.oy()) && (j == 0 || j == g.ny - 1)) {
    const int kz = k < nz - 1 ? k : nz - 1;
    const int ea = j == 0 ? 0 : g.ny - 1, eb = j == 0 ? 1 : g.ny;
    auto V = [&](int jj, int kk) { return double(v[g.at(i, jj, kk)]); };
    const double vb = H * (c.fnm[kz] * (V(eb, kz) + V(ea, kz)) + c.fnp[kz] * (V(eb, kz - 1) + V(ea, kz - 1)));
    const double muf = c.c1f[k] * f.mut[x2] + c.c2f[k];
    if (j == 0) {
      const double vl = vb < 0.0 ? vb : 0.0;
      t = t - c.rdy * muf * (vl * (double(ph[g.at(i, 1, k)]) - ph[g.at(i, 0, k)]));
    } else {
      const double vr = vb > 0.0 ? vb : 0.0;
      t = t - c.rdy * muf * (vr * (double(ph[x]) - ph[x - sj]));
    }
  }
  if (ox && (i == 0 || i == g.nx - 1) && !(i == 0 && k == nz)) {
    const int kz = k < nz ? k : nz - 1;
    const int ea = i == 0 ? 0 : g.nx - 1, eb = i == 0 ? 1 : g.nx;
    auto U = [&](int ii, int kk) { return double(u[g.at(ii, j, kk)]); };
    const double ub = H * (c.fnm[kz] * (U(eb, kz) + U(ea, kz)) + c.fnp[kz] * (U(eb, kz - 1) + U(ea, kz - 1)));
    const double muf = c.c1f[k] * f.mut[x2] + c.c2f[k];
    if (i == 0) {
      const double ul = ub < 0.0 ? ub : 0.0;
      t = t - (f.msftx[x2] / msfty) * c.rdx * muf * (ul * (double(ph[g.at(1, j, k)]) - ph[g.at(0, j, k)]));
    } else {
      const double ur = ub > 0.0 ? ub : 0.0;
      t = t - (f.msftx[x2] / msfty) * c.rdx * muf * (ur * (double(ph[x]) - ph[x - 1]));
    }
  }
  if (sp) t = t + f.ph_tendf[x] / msfty;  // rk_addtend_dry with a LateralBC (large_step.cpp:663)
  f.ph_tend[x] = t;
}

// a / b in the substep kernels: the IEEE division in FP64 (the reference's bits); in FP32-bulk a multiply by the
// hardware reciprocal of b refined by one Newton step (a few FP64 ulps, far below the float storage's rounding),
// which replaces the division's ten-instruction sequence
__device__ __forceinline__ double bdiv(double a, double b) {
  if constexpr (sizeof(real) == 8) {
    return a / b;
  } else {
    const double r = __builtin_amdgcn_rcp(b);
    return a * __builtin_fma(__builtin_fma(-b, r, 1.0), r, r);
  }
}
// The weight of advance_w's implicit Rayleigh damping (damp_opt 3, reference small_step.cpp:393-405) at w level
// k >= 1 of column (i, j): dampmag sin^2(pi/2 (h_k - h_bot) / zdamp) above h_bot = h_top - zdamp, 0 below, the
// heights from the stage's saved geopotential.  It does not change within a stage, so the stage's column kernel
// evaluates it once and the substeps read it.
__device__ __forceinline__ double damp_weight(const Args& a, int i, int j, int k, double dts) {
  const Fields& f = a.f;
  const Geo& g = a.g;
  const double dampmag = dts * a.q.dampcoef;
  const double hdepth = a.q.zdamp;
  const double pi = 3.141592653589793;  // 4 atan(1), folded
  const real ph_top = f.ph_save[g.at(i, j, g.nz)], phs = f.ph_save[g.at(i, j, k)];
  const double htop = bdiv(ph_top + f.phb(i, j, g.nz), kG), hbot = htop - hdepth;
  const double hk = bdiv(phs + f.phb(i, j, k), kG);
  double dampwt = 0;
  if (hk >= hbot) {
    const double sn = mi_libm::sin(bdiv(H * pi * (hk - hbot), hdepth));
    dampwt = dampmag * sn * sn;
  }
  return dampwt;
}

// ----------------------------------------------------------------------------- small_step_prep
// KPrep (reference src/dyn/small_step.cpp:39-93)
MI_POINT_KERNEL k_ss_prep(Range r, Args a, int first) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double mts = f.mub[x2] + f.mu1[x2];
  const double mus_u = first ? f.muu[x2] : H * (mts + f.mub[x2 - 1] + f.mu1[x2 - 1]);
  const double mus_v = first ? f.muv[x2] : H * (mts + f.mub[x2 - g.sj] + f.mu1[x2 - g.sj]);
  if (k == 0) {
    // the map-factor ratios advance_uv reads at every substep: msfux / msfuy, 1 / msfuy (u faces), msfvy / msfvx
    // (v faces)
    if (j < ny) {
      f.rmu[x2] = f.msfux[x2] / f.msfuy[x2];
      f.msfuy_inv[x2] = 1.0 / f.msfuy[x2];
    }
    if (i < nx && j < g.nyv) f.rmv[x2] = f.msfvy[x2] / f.msfvx[x2];
    if (i < nx && j < ny) f.msfty_inv[x2] = 1.0 / f.msfty[x2];  // and the substeps' 1 / msfty
    if (i < nx && j < ny) {
      put<kHMass>(f.muts, g, i, j, 0, mts);
      const double m2 = f.mu2[x2];
      f.mu_save[x2] = m2;
      put<kHMass>(f.mu2, g, i, j, 0, f.mu1[x2] - m2);
    }
    if (j < ny) f.muus[x2] = mus_u;
    if (i < nx && j < g.nyv) f.muvs[x2] = mus_v;
  }
  if (j < ny && k < nz) {
    const double ui = f.u2[x];
    put<kHU>(f.u_save, g, i, j, k * g.sk, ui);
    // the substeps' coupled save (mass x u_save / msfuy), constant within the stage
    f.ucs[x] = bdiv((c.c1h[k] * f.muu[x2] + c.c2h[k]) * double(f.u_save[x]), f.msfuy[x2]);
    put<kHU>(f.u2, g, i, j, k * g.sk, ((c.c1h[k] * mus_u + c.c2h[k]) * f.u1[x] - (c.c1h[k] * f.muu[x2] + c.c2h[k]) * ui) / f.msfuy[x2]);
  }
  if (i < nx && j < g.nyv && k < nz) {
    const double vi = f.v2[x];
    put<kHV>(f.v_save, g, i, j, k * g.sk, vi);
    f.vcs[x] = (c.c1h[k] * f.muv[x2] + c.c2h[k]) * double(f.v_save[x]) * f.msfvx_inv[x2];
    put<kHV>(f.v2, g, i, j, k * g.sk, ((c.c1h[k] * mus_v + c.c2h[k]) * f.v1[x] - (c.c1h[k] * f.muv[x2] + c.c2h[k]) * vi) * f.msfvx_inv[x2]);
  }
  if (i < nx && j < ny) {
    if (k < nz) {
      const double ti = f.t2[x];
      put<kHMass>(f.t_save, g, i, j, k * g.sk, ti);
      f.t2[x] = (c.c1h[k] * mts + c.c2h[k]) * f.t1[x] - (c.c1h[k] * f.mut[x2] + c.c2h[k]) * ti;
      f.c2a[x] = kCpovcv * (f.pb(i, j, k) + f.p[x]) / f.alt[x];
    }
    const double wi = f.w2[x], phi = f.ph2[x];
    f.w_save[x] = wi;
    f.w2[x] = ((c.c1f[k] * mts + c.c2f[k]) * f.w1[x] - (c.c1f[k] * f.mut[x2] + c.c2f[k]) * wi) / f.msfty[x2];
    f.ph_save[x] = phi;
    put<kHMass>(f.ph2, g, i, j, k * g.sk, f.ph1[x] - phi);
    f.ww1[x] = (k == 0 || k == nz) ? 0.0 : f.ww[x];
  }
}

// p_rho_point (small_step.cpp:97-114): al'' and p'' of one point from phi'' at levels k and k+1;
// damp: the smdiv divergence damping of the substeps (KWPRho), else the stage start (KPRho)
__device__ __forceinline__ void p_rho_at(const Args& a, int i, int j, int k, double ph_k, double ph_kp1, double mu2,
                                         double muts, double t2, bool damp) {
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const idx_t x = g.at(i, j, k);
  const double m = c.c1h[k] * muts + c.c2h[k];
  const double al = -((1. / m) * (f.alt[x] * (c.c1h[k] * mu2) + c.rdnw[k] * (ph_kp1 - ph_k)));
  put<kHMass>(f.al, g, i, j, k * g.sk, al);
  double pp = f.c2a[x] * (f.alt[x] * (t2 - c.c1h[k] * mu2 * f.t_save[x]) / (m * (kT0 + f.t_save[x])) - al);
  if (damp) {
    const double ptmp = pp;
    pp = pp + a.q.smdiv * (pp - f.pm1[x]);
    f.pm1[x] = ptmp;
  } else {
    f.pm1[x] = pp;
  }
  put<kHMass>(f.p, g, i, j, k * g.sk, pp);
}

// KPRho (small_step.cpp:116-121) and KCoefW calc_coef_w (small_step.cpp:124-148), dry, one column
// group: p'', al'' and the b, c, a coefficients in parallel, then the alpha / gamma recurrence
MI_COLUMN_KERNEL k_ss_column(Args a, double cof, double dts) {
  constexpr int kKT = KT;
  extern __shared__ double lds[];
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const ColPos cp = col_pos<CI>();
  const int i = cp.i, j = cp.j, nz = g.nz;
  const bool on = i < g.nx;
  double* B = lds;                  // b(k), k = 1..nz
  double* CC = B + (nz + 1) * CI;  // c(k), k = 1..nz-1
  double* AA = CC + (nz + 1) * CI; // a(k)
  if (on) {
    const idx_t x2 = g.at2(i, j);
    const double m = f.mut[x2];
    const double mu2 = f.mu2[x2], muts = f.muts[x2];
    const double lid = a.q.top_lid ? 0.0 : 1.0;
    auto mh = [&](int kk) { return c.c1h[kk] * m + c.c2h[kk]; };
    auto mf = [&](int kk) { return c.c1f[kk] * m + c.c2f[kk]; };
    auto c2a = [&](int kk) { return f.c2a[g.at(i, j, kk)]; };
    for (int k = cp.kt; k <= nz; k += kKT) {
      const idx_t x = g.at(i, j, k);
      if (k < nz) p_rho_at(a, i, j, k, f.ph2[x], f.ph2[x + g.sk], mu2, muts, f.t2[x], false);
      // kern::MoistCoef::cq1 = 1 / (1 + cqw), 1 when dry
      const double cq1 = (a.q.moist && k >= 1 && k < nz) ? 1.0 / (1.0 + f.cqw[x]) : 1.0;
      // for the substeps of the stage: cq1 and advance_w's c2a rdnw / (c1h mut + c2h) at the mass levels
      if (a.q.moist && k >= 1 && k < nz) f.cq1[x] = cq1;
      if (k < nz) f.cwc[x] = bdiv(double(f.c2a[x]) * c.rdnw[k], c.c1h[k] * m + c.c2h[k]);
      double ak;
      if (k <= 1) ak = 0;
      else if (k == nz) ak = -2.0 * cof * (c.rdnw[nz - 1] * c.rdnw[nz - 1]) * c2a(nz - 1) * lid / (mh(nz - 1) * mf(nz - 1));
      else ak = -cq1 * cof * c.rdn[k] * c.rdnw[k - 1] * c2a(k - 1) / (mh(k - 1) * mf(k - 1));
      f.ca[x] = ak;
      LDS(AA, k, cp.ci) = ak;
      if (a.q.damp_opt == 3 && k >= 1) f.dampwt[x] = damp_weight(a, i, j, k, dts);
      if (k >= 1 && k < nz) {
        LDS(B, k, cp.ci) = 1.0 + cq1 * cof * c.rdn[k] * (c.rdnw[k] * c2a(k) / (mh(k) * mf(k)) + c.rdnw[k - 1] * c2a(k - 1) / (mh(k - 1) * mf(k)));
        LDS(CC, k, cp.ci) = -cq1 * cof * c.rdn[k] * c.rdnw[k] * c2a(k) / (mh(k) * mf(k + 1));
      } else if (k == nz) {
        LDS(B, k, cp.ci) = 1.0 + 2.0 * cof * (c.rdnw[nz - 1] * c.rdnw[nz - 1]) * c2a(nz - 1) / (mh(nz - 1) * mf(nz));
      }
    }
  }
  __syncthreads();
  if (on && cp.kt == 0) {
    const idx_t x0 = g.at(i, j, 0), sk = g.sk;
    f.calpha[x0] = 0;
    f.cgamma[x0] = 0;
    double gam = 0.0;
    for (int k0 = 1; k0 < nz; k0 += kChain) {
      double bv[kChain], av[kChain], cv[kChain];
#pragma unroll
      for (int q = 0; q < kChain; ++q)
        if (k0 + q < nz) {
          bv[q] = LDS(B, k0 + q, cp.ci);
          av[q] = LDS(AA, k0 + q, cp.ci);
          cv[q] = LDS(CC, k0 + q, cp.ci);
        }
#pragma unroll
      for (int q = 0; q < kChain; ++q)
        if (k0 + q < nz) {
          const double al = 1.0 / (bv[q] - av[q] * gam);
          f.calpha[x0 + (k0 + q) * sk] = al;
          gam = cv[q] * al;
          f.cgamma[x0 + (k0 + q) * sk] = gam;
        }
    }
    f.calpha[x0 + nz * sk] = 1.0 / (LDS(B, nz, cp.ci) - LDS(AA, nz, cp.ci) * gam);
    f.cgamma[x0 + nz * sk] = 0;
  }
}

// ----------------------------------------------------------------------------- advance_uv
// FP32-bulk: the same update with the bulk arithmetic in FP32 (one multiply by a rounded reciprocal per division);
// the differences of the full geopotential php and of the base pressure pb are taken in FP64 first
template <bool OPEN>
__device__ __forceinline__ void uv_fp32(const Args& a, int i, int j, int k, double dts_d) {
  using F = float;
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j), sk = g.sk;
  const F dts = F(dts_d);
  auto dpn = [&](idx_t d, int kk) -> F {
    auto ps = [&](int m) { const idx_t y = g.at(i, j, m); return F(f.p[y]) + F(f.p[y - d]); };
    if (kk == 0) return F(0.5) * (F(c.cf1) * ps(0) + F(c.cf2) * ps(1) + F(c.cf3) * ps(2));
    if (kk == nz) return a.q.top_lid ? F(0.5) * (F(c.cf1) * ps(nz - 1) + F(c.cf2) * ps(nz - 2) + F(c.cf3) * ps(nz - 3)) : F(0);
    return F(0.5) * (F(c.fnm[kk]) * ps(kk) + F(c.fnp[kk]) * ps(kk - 1));
  };
  auto dp = [&](idx_t d, int di, int dj, double mface, double rd, double rm) -> F {
    const F c1 = F(c.c1h[k]);
    F dd = F(rm * 0.5 * rd) * (c1 * F(mface) + F(c.c2h[k])) *
           (((F(f.ph2[x + sk]) - F(f.ph2[x + sk - d])) + (F(f.ph2[x]) - F(f.ph2[x - d]))) +
            (F(f.alt[x]) + F(f.alt[x - d])) * (F(f.p[x]) - F(f.p[x - d])) +
            (F(f.al[x]) + F(f.al[x - d])) * F(f.pb(i, j, k) - f.pb(i - di, j - dj, k)));
    dd += F(rm * rd) * F(f.php[x] - f.php[x - d]) *
          (F(c.rdnw[k]) * (dpn(d, k + 1) - dpn(d, k)) - F(0.5) * c1 * (F(f.mu2[x2 - d]) + F(f.mu2[x2])));
    return dd;
  };
  const int nx = g.nx, ny = g.ny;
  const bool sp = g.spec();
  const bool uface = sp && (i <= 0 || i >= nx || j == 0 || j == ny - 1);
  const bool vface = sp && (i == 0 || i == nx - 1 || j <= 0 || j >= ny);
  const int ui0 = g.bcx != MI_BC_PERIODIC ? 1 : 0;
  if ((uface || ((OPEN && g.ox()) && (i == 0 || i == nx))) && j < ny) {
    f.u2[x] = real(F(f.u2[x]) + dts * F(f.ru_tend[x]));
  } else if (!uface && i >= ui0 && i < nx && j < ny) {
    const F mudf_x = F(-a.q.emdx * (f.mudf[x2] - f.mudf[x2 - 1]) * f.msfuy_inv[x2]);
    const F cq = a.q.moist ? F(f.cqu[x]) : F(1);
    f.u2[x] = real(F(f.u2[x]) + dts * F(f.ru_tend[x]) - dts * cq * dp(1, 1, 0, f.muu[x2], c.rdx, f.rmu[x2]) +
                   F(c.c1h[k]) * mudf_x);
  }
  const int vj0 = (g.bcy != MI_BC_PERIODIC && !g.two_d) ? 1 : 0;
  if ((vface || ((OPEN && g.oy()) && (j == 0 || j == ny))) && i < nx) {
    f.v2[x] = real(F(f.v2[x]) + dts * F(f.rv_tend[x]));
  } else if (!vface && !g.two_d && j >= vj0 && j < ny && i < nx) {
    const idx_t sj = g.sj;
    const F mudf_y = F(-a.q.emdy * (f.mudf[x2] - f.mudf[x2 - sj]) * f.msfvx_inv[x2]);
    const F cq = a.q.moist ? F(f.cqv[x]) : F(1);
    f.v2[x] = real(F(f.v2[x]) + dts * F(f.rv_tend[x]) - dts * cq * dp(sj, 0, 1, f.muv[x2], c.rdy, f.rmv[x2]) +
                   F(c.c1h[k]) * mudf_y);
  }
}

// KUV (small_step.cpp:151-190)
template <bool OPEN>
__device__ __forceinline__ void uv_at(const Args& a, int i, int j, int k, double dts) {
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j), sk = g.sk;
  if constexpr (sizeof(real) == 4) {
    uv_fp32<OPEN>(a, i, j, k, dts);
    return;
  }
  auto dpn = [&](idx_t d, int kk) {
    auto ps = [&](int m) { const idx_t y = g.at(i, j, m); return f.p[y] + f.p[y - d]; };
    if (kk == 0) return H * (c.cf1 * ps(0) + c.cf2 * ps(1) + c.cf3 * ps(2));
    if (kk == nz) return a.q.top_lid ? H * (c.cf1 * ps(nz - 1) + c.cf2 * ps(nz - 2) + c.cf3 * ps(nz - 3)) : 0.0;
    return H * (c.fnm[kk] * ps(kk) + c.fnp[kk] * ps(kk - 1));
  };
  auto dp = [&](idx_t d, int di, int dj, double mface, double rd, double rm) {
    double dd = rm * H * rd * (c.c1h[k] * mface + c.c2h[k]) *
                (((f.ph2[x + sk] - f.ph2[x + sk - d]) + (f.ph2[x] - f.ph2[x - d])) + (f.alt[x] + f.alt[x - d]) * (f.p[x] - f.p[x - d]) +
                 (f.al[x] + f.al[x - d]) * (f.pb(i, j, k) - f.pb(i - di, j - dj, k)));
    dd += rm * rd * (f.php[x] - f.php[x - d]) *
          (c.rdnw[k] * (dpn(d, k + 1) - dpn(d, k)) - H * (c.c1h[k] * f.mu2[x2 - d] + c.c1h[k] * f.mu2[x2]));
    return dd;
  };
  // Edge (small_step.cpp:39-45): with specified boundaries the spec zone's faces are not advanced here but by
  // spec_bdyupdate, u += dts ru_tend (LateralBC::update_uv, lateral_bc.cpp:356-364), fused in; the launch then
  // covers the east face i = nx and the north face j = ny too
  const int nx = g.nx, ny = g.ny;
  // Open boundaries (small_step.cpp:175-184): the boundary faces take the large-step tendency only
  const bool sp = g.spec();
  const bool uface = sp && (i <= 0 || i >= nx || j == 0 || j == ny - 1);
  const bool vface = sp && (i == 0 || i == nx - 1 || j <= 0 || j >= ny);
  const int ui0 = g.bcx != MI_BC_PERIODIC ? 1 : 0;
  if ((uface || ((OPEN && g.ox()) && (i == 0 || i == nx))) && j < ny) {
    put<kHU>(f.u2, g, i, j, k * sk, f.u2[x] + dts * f.ru_tend[x]);
  } else if (!uface && i >= ui0 && i < nx && j < ny) {
    const double mudf_x = -a.q.emdx * (f.mudf[x2] - f.mudf[x2 - 1]) / f.msfuy[x2];
    put<kHU>(f.u2, g, i, j, k * sk, f.u2[x] + dts * f.ru_tend[x] - dts * (a.q.moist ? f.cqu[x] : 1.0) * dp(1, 1, 0, f.muu[x2], c.rdx, f.rmu[x2]) + c.c1h[k] * mudf_x);
  }
  const int vj0 = (g.bcy != MI_BC_PERIODIC && !g.two_d) ? 1 : 0;
  if ((vface || ((OPEN && g.oy()) && (j == 0 || j == ny))) && i < nx) {
    put<kHV>(f.v2, g, i, j, k * sk, f.v2[x] + dts * f.rv_tend[x]);
  } else if (!vface && !g.two_d && j >= vj0 && j < ny && i < nx) {
    const idx_t sj = g.sj;
    const double mudf_y = -a.q.emdy * (f.mudf[x2] - f.mudf[x2 - sj]) * f.msfvx_inv[x2];
    put<kHV>(f.v2, g, i, j, k * sk, f.v2[x] + dts * f.rv_tend[x] - dts * (a.q.moist ? f.cqv[x] : 1.0) * dp(sj, 0, 1, f.muv[x2], c.rdy, f.rmv[x2]) + c.c1h[k] * mudf_y);
  }
}
template <bool OPEN>
MI_POINT_KERNEL k_uv(Range r, Args a, double dts) {
  MI_POINT_INDEX(r)
  uv_at<OPEN>(a, i, j, k, dts);
}

// ----------------------------------------------------------------------------- one acoustic substep's column work
// FP32-bulk with 64 threads per column: the column recurrences run as wave-wide scans, one wavefront per column
// (wave w of the block takes column w, lane k level k).  The sums then associate differently and the Thomas
// sweeps become compositions of the affine maps y_k = p_k + q_k y_(k-1): the same solution in another rounding,
// within FP32-bulk's error bar.  FP64 keeps the reference's serial order (bit-identical).
constexpr bool kScanChains = sizeof(real) == 4;
// inclusive prefix sum over the lanes of a wavefront
__device__ __forceinline__ double wave_prefix_sum(double v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}
// y_k = p_k + q_k y_(k-1) composed from lane 0 up (up = true) or y_k = p_k + q_k y_(k+1) from lane 63 down
template <bool UP>
__device__ __forceinline__ double wave_affine_scan(double p, double q, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double pp = UP ? __shfl_up(p, d, 64) : __shfl_down(p, d, 64);
    const double qq = UP ? __shfl_up(q, d, 64) : __shfl_down(q, d, 64);
    if (UP ? lane >= d : lane + d < 64) {
      p = p + q * pp;
      q = q * qq;
    }
  }
  return p;
}
// advance_mu_t + sumflux and advance_w + calc_p_rho of one substep in one kernel (reference
// SmallStep::run_stage, small_step.cpp:548-593).  Between KMuTPoint and KWRhs the reference refreshes
// the halos of mu, mudf and muts; none of the advance_w kernels reads a halo point (each reads its own
// column only), so the refresh moves after this kernel.
//
// One thread per level (KT >= nz + 1).  Every global load a thread needs is issued in one burst at the
// start, so a block waits on memory once; the phases after it exchange values through LDS only.  The
// recurrences of the column keep their order and operands: the per-level terms they add are computed in
// parallel first (the same products and quotients the reference forms inside its loops), then one thread
// per column chains the additions.
//   1  KMuTDiv dvdxi (small_step.cpp:194-207) and the dnw*dvdxi terms of dmdt
//   2  KMuTCol: dmdt, mu, mudf, muts, muave (small_step.cpp:218-230), one thread
//   3  KMuTCol's omega'' terms dnw*(c1h*dmdt + dvdxi + c1h*mu_tend)/msfty, parallel
//   4  the running omega'' integral, one thread
//   5  omega'' = integral - ww1; KMuTCol's ww_m sums; KMuTPoint (small_step.cpp:250-276): ru_m, rv_m,
//      theta''; KWRhs (small_step.cpp:312-328): explicit phi'' and t_2ave
//   6  KWExpl (small_step.cpp:331-378): explicit w''
//   7  KWSolve's tridiagonal sweeps (small_step.cpp:386-387), one thread
//   8  KWSolve's damp_opt 3 damping and phi'' (small_step.cpp:388-406), parallel over levels
//   9  KWPRho (small_step.cpp:409-414)
// With specified boundaries the spec zone's columns (Edge::mass, small_step.cpp:39-45) are skipped here and
// advanced by k_acoustic_ring.
template <int CI, int KT>
__device__ __forceinline__ void acoustic_cols(const Args& a, const ColPos cp, double* lds, double dts, int it, int ns) {
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int i = cp.i, j = cp.j, nz = g.nz, ci = cp.ci, k = cp.kt;
  // with specified boundaries the spec zone's columns are left to k_acoustic_ring
  const bool ring = g.bcx == MI_BC_SPECIFIED && (i == 0 || i == g.nx - 1 || j == 0 || j == g.ny - 1);
  const bool on = i < g.nx && k <= nz && !ring;
  const int L = (nz + 1) * CI;
  double* DV = lds;     // dnw*dvdxi(k), then dvdxi(k)
  double* TK = DV + L;  // the omega'' terms, then the integral
  double* WW = TK + L;  // omega''(k)
  double* R = WW + L;   // explicit phi''(k)
  double* T2A = R + L;  // t_2ave(k)
  double* W = T2A + L;  // w''(k)
  double* AA = W + L;   // calc_coef_w a(k), alpha(k), gamma(k)
  double* AL = AA + L;
  double* GA = AL + L;
  double* PH = GA + L;  // phi'' after the solve
  double* S2 = PH + L;  // per column: dmdt, mu'' new, muts, muave
  const bool first = it == 1, last = it == ns;
  const double rn = double(ns);
  const double epssm = a.q.epssm;
  const idx_t sk = g.sk, sj = g.sj;
  const idx_t x2 = g.at2(i, j), x = g.at(i, j, k);
  const bool mk = k < nz, ge1 = k >= 1;  // a mass level; above the surface

  // ---------------------------------------------------------------- the load burst
  // the values of the 3-D bulk fields stay in their storage type (half the registers in FP32-bulk); every
  // expression below converts them to double before it combines two of them
  // ucs, vcs: the stage's coupled saves (mass x u_save / msfuy, mass x v_save / msfvx) of the two faces
  real u0 = 0, u1 = 0, ucs0 = 0, ucs1 = 0, v0 = 0, v1 = 0, vcs0 = 0, vcs1 = 0;
  real ww1 = 0, wwm = 0, rum = 0, rvm = 0;
  double ca = 0, cal = 0, cga = 0;
  real t2 = 0, ttend = 0, ts = 0, tsxp = 0, tsxm = 0, tsyp = 0, tsym = 0, tskm = 0, tskp = 0;
  real phtend = 0, w2 = 0, phs = 0, phs_m = 0, phs_p = 0, ph = 0, ph_m = 0, ph_p = 0, rwt = 0;
  real c2a = 0, c2a_m = 0, alt = 0, alt_m = 0, wsv = 0, pm1 = 0;
  real cw = 0, cw_m = 0, cq1 = 1;  // the stage's c2a rdnw / (c1h mut + c2h) at k and k - 1, 1 / (1 + cqw)
  double msfty = 1, msftx = 1, msft_inv = 1, mut = 0;
  double mu_old = 0, mtend = 0;
  real dwt = 0;
  if (on) {
    msfty = f.msfty[x2];
    msftx = f.msftx[x2];
    msft_inv = f.msfty_inv[x2];
    mut = f.mut[x2];
    if (k == 0) {
      mu_old = f.mu2[x2];
      mtend = f.mu_tend[x2];
    }
    ww1 = f.ww1[x];
    if (!first) wwm = f.ww_m[x];
    ca = f.ca[x];
    cal = f.calpha[x];
    cga = f.cgamma[x];
    phtend = f.ph_tend[x];
    w2 = f.w2[x];
    phs = f.ph_save[x];
    ph = f.ph2[x];
    if (ge1) {
      phs_m = f.ph_save[x - sk];
      ph_m = f.ph2[x - sk];
      rwt = f.rw_tend[x];
      c2a_m = f.c2a[x - sk];
      alt_m = f.alt[x - sk];
      cw_m = f.cwc[x - sk];
      if (mk && a.q.moist) cq1 = f.cq1[x];
      if (a.q.damp_opt == 3) {
        wsv = f.w_save[x];
        dwt = f.dampwt[x];
      }
    }
    if (mk) {
      u0 = f.u2[x];
      u1 = f.u2[x + 1];
      ucs0 = f.ucs[x];
      ucs1 = f.ucs[x + 1];
      cw = f.cwc[x];
      if (!g.two_d) {
        v0 = f.v2[x];
        v1 = f.v2[x + sj];
        vcs0 = f.vcs[x];
        vcs1 = f.vcs[x + sj];
      }
      if (!first) {
        rum = f.ru_m[x];
        if (!g.two_d) rvm = f.rv_m[x];
      }
      t2 = f.t2[x];
      ttend = f.t_tend[x];
      ts = f.t_save[x];
      tsxp = f.t_save[x + 1];
      tsxm = f.t_save[x - 1];
      tsyp = f.t_save[x + sj];
      tsym = f.t_save[x - sj];
      if (ge1) tskm = f.t_save[x - sk];
      if (k + 1 < nz) tskp = f.t_save[x + sk];
      phs_p = f.ph_save[x + sk];
      ph_p = f.ph2[x + sk];
      c2a = f.c2a[x];
      alt = f.alt[x];
      pm1 = f.pm1[x];
    }
  }

  // ---------------------------------------------------------------- 1: dvdxi
  double dv = 0;
  if (on) {
    LDS(AA, k, ci) = ca;
    LDS(AL, k, ci) = cal;
    LDS(GA, k, ci) = cga;
    if (mk) {
      const double fx = (u1 + double(ucs1)) - (u0 + double(ucs0));
      double d = c.rdx * fx;
      if (!g.two_d) d += c.rdy * ((v1 + double(vcs1)) - (v0 + double(vcs0)));
      dv = msftx * msfty * d;
      LDS(DV, k, ci) = c.dnw[k] * dv;
    }
  }
  __syncthreads();
  // ---------------------------------------------------------------- 2: the column mass
  // the scan layout: wave sw = column, lane sl = level; sw_on: that column is stepped here
  const int sw = int(threadIdx.x) / 64, sl = int(threadIdx.x) % 64;
  const int sw_i = i - ci + sw;
  const bool sw_on = kScanChains && KT == 64 && sw_i < g.nx &&
                     !(g.bcx == MI_BC_SPECIFIED && (sw_i == 0 || sw_i == g.nx - 1 || j == 0 || j == g.ny - 1));
  if constexpr (kScanChains && KT == 64) {
    double v = (sw_on && sl < nz) ? LDS(DV, sl, sw) : 0.0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    if (sw_on && sl == 0) S2[sw] = v;
    __syncthreads();
  }
  if (on && k == 0) {
    double dmdt = 0;
    if constexpr (kScanChains && KT == 64) {
      dmdt = S2[ci];
    } else {
      for (int kk = 0; kk < nz; ++kk) dmdt += LDS(DV, kk, ci);
    }
    const double mu_new = mu_old + dts * (dmdt + mtend);
    const double muts = mut + mu_new;
    const double muave = H * ((1.0 + epssm) * mu_new + (1.0 - epssm) * mu_old);
    put<kHMass>(f.mu2, g, i, j, 0, mu_new);
    put<kHMass>(f.mudf, g, i, j, 0, dmdt + mtend);
    put<kHMass>(f.muts, g, i, j, 0, muts);
    f.muave[x2] = muave;
    S2[ci] = dmdt;
    S2[CI + ci] = mu_new;
    S2[2 * CI + ci] = muts;
    S2[3 * CI + ci] = muave;
    S2[4 * CI + ci] = mtend;
  }
  __syncthreads();
  const double dmdt = S2[ci], mu_new = S2[CI + ci], muts = S2[2 * CI + ci], muave = S2[3 * CI + ci];
  // ---------------------------------------------------------------- 3: the omega'' terms
  if (on && k + 1 < nz) {
    const double mt = S2[4 * CI + ci];
    const double tk = c.dnw[k] * (c.c1h[k] * dmdt + dv + c.c1h[k] * mt);
    // FP64: the reference's division; FP32-bulk: the multiply by the stage's 1 / msfty
    if constexpr (sizeof(real) == 8) LDS(TK, k, ci) = tk / msfty;
    else LDS(TK, k, ci) = tk * msft_inv;
  }
  __syncthreads();
  // ---------------------------------------------------------------- 4: the running integral
  if constexpr (kScanChains && KT == 64) {
    const bool in = sw_on && sl + 1 < nz;
    const double v = wave_prefix_sum(in ? LDS(TK, sl, sw) : 0.0, sl);
    if (in) LDS(TK, sl, sw) = -v;  // integral up to level sl + 1
  } else if (on && k == 0) {
    double ww_full = 0;
    for (int kk = 0; kk + 1 < nz; ++kk) {
      ww_full = ww_full - LDS(TK, kk, ci);
      LDS(TK, kk, ci) = ww_full;  // integral up to level kk + 1
    }
  }
  __syncthreads();
  // ---------------------------------------------------------------- 5: omega'', theta'', explicit phi''
  double wwk = 0;
  if (on) {
    wwk = (k >= 1 && k < nz) ? LDS(TK, k - 1, ci) - ww1 : 0.0;
    LDS(WW, k, ci) = wwk;
  }
  __syncthreads();
  double t2new = 0;
  if (on) {
    f.ww[x] = wwk;
    {
      const double sw = (first ? 0.0 : wwm) + wwk;
      put<kHMass>(f.ww_m, g, i, j, k * sk, last ? bdiv(sw, rn) + ww1 : sw);
    }
    if (mk) {
      // wdtn(k) and wdtn(k+1): omega'' times theta at the w level (0 at the bottom and the top)
      const double wdtn_k = (k == 0) ? 0.0 : wwk * (c.fnm[k] * ts + c.fnp[k] * tskm);
      const double wdtn_kp = (k + 1 == nz) ? 0.0 : LDS(WW, k + 1, ci) * (c.fnm[k + 1] * tskp + c.fnp[k + 1] * ts);
      const double su = (first ? 0.0 : rum) + u0;
      put<kHU>(f.ru_m, g, i, j, k * sk, last ? bdiv(su, rn) + double(ucs0) : su);
      if (!g.two_d) {
        const double sv = (first ? 0.0 : rvm) + v0;
        put<kHV>(f.rv_m, g, i, j, k * sk, last ? bdiv(sv, rn) + double(vcs0) : sv);
      }
      double hflux = H * c.rdx * (double(u1) * (double(tsxp) + ts) - double(u0) * (double(ts) + tsxm));
      if (!g.two_d) hflux += H * c.rdy * (double(v1) * (double(tsyp) + ts) - double(v0) * (double(ts) + tsym));
      const double tn = t2 + msfty * dts * ttend;
      t2new = tn - dts * msfty * (msftx * hflux + c.rdnw[k] * (wdtn_kp - wdtn_k));
      f.t2[x] = t2new;
      const double tav = H * ((1.0 + epssm) * t2new + (1.0 - epssm) * t2);
      LDS(T2A, k, ci) = bdiv(tav + c.c1h[k] * muave * kT0, (c.c1h[k] * muts + c.c2h[k]) * (kT0 + ts));
    }
    double Rk;
    if (k == 0 || (a.q.top_lid && k == nz)) {
      Rk = 0;
    } else {
      // wdwn(m): omega'' at the mass level times the layer's geopotential difference
      const double wdwn_k = H * (wwk + LDS(WW, k - 1, ci)) * c.rdnw[k - 1] * (double(phs) - phs_m + f.phb(i, j, k) - f.phb(i, j, k - 1));
      double rr = dts * (phtend + H * kG * (1.0 - epssm) * w2);
      if (mk) {
        const double wdwn_kp = H * (LDS(WW, k + 1, ci) + wwk) * c.rdnw[k] * (double(phs_p) - phs + f.phb(i, j, k + 1) - f.phb(i, j, k));
        rr -= dts * (c.fnm[k] * wdwn_kp + c.fnp[k] * wdwn_k);
      }
      Rk = ph + bdiv(msfty * rr, c.c1f[k] * mut + c.c2f[k]);
    }
    LDS(R, k, ci) = Rk;
  }
  __syncthreads();
  // ---------------------------------------------------------------- 6: explicit w''
  if (on) {
    const double m = mut, mav = muave;
    const double ope = 1.0 + epssm, ome = 1.0 - epssm;
    double wk;
    if (k == 0) {
      // the surface w'': 0 on flat ground; with terrain, from the small-step u'', v'' and the slope of ht
      // (small_step.cpp:339-353)
      wk = 0;
      if (a.q.terrain) {
        const auto* ht = f.ht;
        const auto* us = f.u2;
        const auto* vs = f.v2;
        wk = msftx * H * c.rdx *
             ((ht[x2 + 1] - ht[x2]) * (c.cf1 * us[x + 1] + c.cf2 * us[x + 1 + sk] + c.cf3 * us[x + 1 + 2 * sk]) +
              (ht[x2] - ht[x2 - 1]) * (c.cf1 * us[x] + c.cf2 * us[x + sk] + c.cf3 * us[x + 2 * sk]));
        if (!g.two_d)
          wk = msfty * H * c.rdy *
                   ((ht[x2 + sj] - ht[x2]) * (c.cf1 * vs[x + sj] + c.cf2 * vs[x + sj + sk] + c.cf3 * vs[x + sj + 2 * sk]) +
                    (ht[x2] - ht[x2 - sj]) * (c.cf1 * vs[x] + c.cf2 * vs[x + sk] + c.cf3 * vs[x + 2 * sk])) +
               wk;
      }
    } else {
      const double r_k = LDS(R, k, ci), r_km1 = LDS(R, k - 1, ci);
      if (mk) {
        const double r_kp1 = LDS(R, k + 1, ci);
        wk = w2 + dts * rwt +
             msft_inv * (a.q.moist ? double(cq1) : 1.0) *
                 (H * dts * kG * c.rdn[k] *
                  (double(cw) * (ope * (r_kp1 - r_k) + ome * (double(ph_p) - ph)) -
                   double(cw_m) * (ope * (r_k - r_km1) + ome * (double(ph) - ph_m)))) +
             dts * kG * msft_inv * (c.rdn[k] * (double(c2a) * alt * LDS(T2A, k, ci) - double(c2a_m) * alt_m * LDS(T2A, k - 1, ci)) - c.c1f[k] * mav);
      } else {
        const int km = nz - 1;
        wk = w2 + dts * rwt +
             msft_inv * (bdiv(-H * dts * kG, c.c1h[km] * m + c.c2h[km]) * (c.rdnw[km] * c.rdnw[km]) * 2.0 * c2a_m * (ope * (r_k - r_km1) + ome * (double(ph) - ph_m)) -
                         dts * kG * (2.0 * c.rdnw[km] * c2a_m * alt_m * LDS(T2A, km, ci) + c.c1f[k] * mav));
        if (a.q.top_lid) wk = 0;
      }
    }
    LDS(W, k, ci) = wk;
  }
  __syncthreads();
  // ---------------------------------------------------------------- 7: tridiagonal sweeps
  if constexpr (kScanChains && KT == 64) {
    // forward: y_0 = w_0, y_k = (w_k - a_k y_(k-1)) alpha_k; back: w_nz = y_nz, w_k = y_k - gamma_k w_(k+1)
    double p = 0.0, q = 1.0;
    if (sw_on && sl == 0) {
      p = LDS(W, 0, sw);
      q = 0.0;
    } else if (sw_on && sl <= nz) {
      const double al = LDS(AL, sl, sw);
      p = LDS(W, sl, sw) * al;
      q = -(LDS(AA, sl, sw) * al);
    }
    const double y = wave_affine_scan<true>(p, q, sl);
    double qb = 1.0, pb = 0.0;
    if (sw_on && sl <= nz) {
      pb = y;
      qb = (sl >= 1 && sl < nz) ? -LDS(GA, sl, sw) : 0.0;
    }
    const double wsol = wave_affine_scan<false>(pb, qb, sl);
    if (sw_on && sl <= nz) LDS(W, sl, sw) = wsol;
  } else if (on && k == 0) {
    double wprev = LDS(W, 0, ci);
    for (int kk = 1; kk <= nz; ++kk) {
      wprev = (LDS(W, kk, ci) - LDS(AA, kk, ci) * wprev) * LDS(AL, kk, ci);
      LDS(W, kk, ci) = wprev;
    }
    double wnext = LDS(W, nz, ci);
    for (int kk = nz - 1; kk >= 1; --kk) {
      wnext = LDS(W, kk, ci) - LDS(GA, kk, ci) * wnext;
      LDS(W, kk, ci) = wnext;
    }
  }
  __syncthreads();
  // ---------------------------------------------------------------- 8: damping near the top, phi''
  if (on) {
    double wk = LDS(W, k, ci);
    double phk = ph;  // the surface phi'' is not solved for
    if (k >= 1) {
      const double dampmag = a.q.damp_opt == 3 ? dts * a.q.dampcoef : 0.0;
      if (dampmag != 0.0) {  // the stage's weight (damp_weight)
        const double dampwt = dwt;
        wk = bdiv(wk - dampwt * (c.c1f[k] * mut + c.c2f[k]) * wsv, 1.0 + dampwt);
      }
      phk = LDS(R, k, ci) + bdiv(msfty * H * dts * kG * (1.0 + epssm) * wk, c.c1f[k] * muts + c.c2f[k]);
      put<kHMass>(f.ph2, g, i, j, k * sk, phk);
    }
    f.w2[x] = wk;
    LDS(PH, k, ci) = phk;
  }
  __syncthreads();
  // ---------------------------------------------------------------- 9: p'', al'' with divergence damping
  if (on && mk) {
    const double m = c.c1h[k] * muts + c.c2h[k];
    const double al = -(bdiv(1., m) * (alt * (c.c1h[k] * mu_new) + c.rdnw[k] * (LDS(PH, k + 1, ci) - LDS(PH, k, ci))));
    put<kHMass>(f.al, g, i, j, k * sk, al);
    double pp = c2a * (bdiv(alt * (t2new - c.c1h[k] * mu_new * ts), m * (kT0 + ts)) - al);
    const double ptmp = pp;
    pp = pp + a.q.smdiv * (pp - pm1);
    f.pm1[x] = ptmp;
    put<kHMass>(f.p, g, i, j, k * sk, pp);
  }
}
MI_COLUMN_KERNEL k_acoustic(Args a, double dts, int it, int ns) {
  extern __shared__ double lds[];
  acoustic_cols<CI, KT>(a, col_pos<CI>(), lds, dts, it, ns);
}

// ----------------------------------------------------------------------------- one acoustic substep, k-march
// The substep's column work (the same phases and the same expressions as k_acoustic) with one thread per column: a
// wavefront holds 64 consecutive columns of one row and marches up and down them, so every load and store of the
// wavefront is one i-contiguous run of a row and no lane idles on a level its column does not have.  No LDS and no
// barriers: a column's recurrences run in its own thread in the reference's order (k upward; the sweeps up, then
// down), its vertical neighbours slide through registers.  Three passes over the column:
//   1 (up)    the mass divergence of every level and its column sum: mu'', mudf, muts, muave
//   2 (up)    omega'' (the running integral), ww_m, ru_m, rv_m, theta'', t_2ave, the explicit phi'' R and w'' and the
//             forward sweep of the tridiagonal solve, whose values y and R go to two FP64 scratch arrays
//   3 (down)  the back-substitution, damp_opt 3, phi'', then p'' and al'' with divergence damping
// The divergence is recomputed in pass 2 from the same loads (the same bits as pass 1's), not stored.  Every
// expression is k_acoustic's, operand for operand, in FP64 arithmetic, so the FP64 mode gives the reference's
// bits; FP32-bulk rounds only its stores.
__global__ void __launch_bounds__(64) k_acoustic_march(Args a, double dts, int it, int ns) {
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const Blk b = xcd_block();
  const int i = b.x * 64 + int(threadIdx.x), j = b.y, nz = g.nz;
  if (i >= g.nx) return;
  if (g.bcx == MI_BC_SPECIFIED && (i == 0 || i == g.nx - 1 || j == 0 || j == g.ny - 1)) return;  // k_acoustic_ring
  const bool first = it == 1, last = it == ns, two_d = g.two_d;
  const double rn = double(ns);
  const double epssm = a.q.epssm;
  const idx_t sk = g.sk, sj = g.sj;
  const idx_t x2 = g.at2(i, j), x0 = g.at(i, j, 0);
  // the column's 2-D values
  const double msfty = f.msfty[x2], msftx = f.msftx[x2], mut = f.mut[x2];
  const double muu0 = f.muu[x2], muu1 = f.muu[x2 + 1], msfuy0 = f.msfuy[x2], msfuy1 = f.msfuy[x2 + 1];
  double muv0 = 0, muv1 = 0, mvi0 = 1, mvi1 = 1;
  if (!two_d) {
    muv0 = f.muv[x2];
    muv1 = f.muv[x2 + sj];
    mvi0 = f.msfvx_inv[x2];
    mvi1 = f.msfvx_inv[x2 + sj];
  }
  // the mass divergence dvdxi of level k (KMuTDiv, small_step.cpp:194-207) and the level's u'', v'', saves
  struct Lev {
    real u0, u1, us0, v0, v1, vs0;
  };
  auto divergence = [&](int k, idx_t x, Lev& l) -> double {
    l.u0 = f.u2[x];
    l.u1 = f.u2[x + 1];
    l.us0 = f.u_save[x];
    const real us1 = f.u_save[x + 1];
    const double fx = (l.u1 + (c.c1h[k] * muu1 + c.c2h[k]) * us1 / msfuy1) - (l.u0 + (c.c1h[k] * muu0 + c.c2h[k]) * l.us0 / msfuy0);
    double d = c.rdx * fx;
    l.v0 = l.v1 = l.vs0 = 0;
    if (!two_d) {
      l.v0 = f.v2[x];
      l.v1 = f.v2[x + sj];
      l.vs0 = f.v_save[x];
      const real vs1 = f.v_save[x + sj];
      d += c.rdy * ((l.v1 + (c.c1h[k] * muv1 + c.c2h[k]) * vs1 * mvi1) - (l.v0 + (c.c1h[k] * muv0 + c.c2h[k]) * l.vs0 * mvi0));
    }
    return msftx * msfty * d;
  };
  // ---------------------------------------------------------------- pass 1: dmdt and the column masses
  double dmdt = 0;
  real su[3][4] = {};  // u'' (i, i+1) and v'' (j, j+1) of levels 0..2, for the surface w'' over terrain
  for (int k = 0; k < nz; ++k) {
    Lev l;
    const double dv = divergence(k, x0 + k * sk, l);
    dmdt += c.dnw[k] * dv;
    if (k < 3) {
      su[k][0] = l.u0;
      su[k][1] = l.u1;
      su[k][2] = l.v0;
      su[k][3] = l.v1;
    }
  }
  const double mu_old = f.mu2[x2], mtend = f.mu_tend[x2];
  const double mu_new = mu_old + dts * (dmdt + mtend);
  const double muts = mut + mu_new;
  const double muave = H * ((1.0 + epssm) * mu_new + (1.0 - epssm) * mu_old);
  put<kHMass>(f.mu2, g, i, j, 0, mu_new);
  put<kHMass>(f.mudf, g, i, j, 0, dmdt + mtend);
  put<kHMass>(f.muts, g, i, j, 0, muts);
  f.muave[x2] = muave;
  // ---------------------------------------------------------------- pass 2: omega'', theta'', explicit phi'' and w''
  const double ope = 1.0 + epssm, ome = 1.0 - epssm;
  const double msft_inv = 1.0 / msfty;
  // the surface w'' (small_step.cpp:339-353): 0 on flat ground; with terrain from u'', v'' and the slope of ht
  double w_expl0 = 0;
  if (a.q.terrain) {
    const double* ht = f.ht;
    w_expl0 = msftx * H * c.rdx *
              ((ht[x2 + 1] - ht[x2]) * (c.cf1 * su[0][1] + c.cf2 * su[1][1] + c.cf3 * su[2][1]) +
               (ht[x2] - ht[x2 - 1]) * (c.cf1 * su[0][0] + c.cf2 * su[1][0] + c.cf3 * su[2][0]));
    if (!two_d)
      w_expl0 = msfty * H * c.rdy *
                    ((ht[x2 + sj] - ht[x2]) * (c.cf1 * su[0][3] + c.cf2 * su[1][3] + c.cf3 * su[2][3]) +
                     (ht[x2] - ht[x2 - sj]) * (c.cf1 * su[0][2] + c.cf2 * su[1][2] + c.cf3 * su[2][2])) +
                w_expl0;
  }
  // sliding windows (levels n-2 .. n+1 of the level n being worked on)
  double ww_full = 0;                     // the running omega'' integral (KMuTCol's ww_full)
  double ww_m1 = 0, ww_0 = 0, ww_p1 = 0;  // omega'' at w levels n-1, n, n+1
  real ts_m1 = 0, ts_0 = f.t_save[x0], ts_p1 = 0;
  real phs_m1 = 0, phs_0 = f.ph_save[x0], phs_p1 = 0;
  double phb_m1 = 0, phb_0 = f.phb(i, j, 0), phb_p1 = 0;
  real ph_m2 = 0, ph_m1 = 0, ph_0 = f.ph2[x0];
  double R_m2 = 0, R_m1 = 0;
  double t2a_m2 = 0, t2a_m1 = 0;
  real c2a_m2 = 0, c2a_m1 = 0, alt_m2 = 0, alt_m1 = 0;
  double y_prev = 0;
  // the explicit w'' at w level k (KWExpl, small_step.cpp:331-378)
  auto w_explicit = [&](int k, double r_k, double r_km1, double r_kp1, real ph, real ph_m, real ph_p, double t2a_k,
                        double t2a_km, real c2a, real c2a_m, real alt, real alt_m) -> double {
    const idx_t xk = x0 + k * sk;
    const double m = mut, mav = muave;
    const real w2 = f.w2[xk], rwt = f.rw_tend[xk];
    if (k < nz) {
      return w2 + dts * rwt +
             msft_inv * (a.q.moist ? 1.0 / (1.0 + f.cqw[xk]) : 1.0) *
                 (H * dts * kG * c.rdn[k] *
                  (c2a * c.rdnw[k] / (c.c1h[k] * m + c.c2h[k]) * (ope * (r_kp1 - r_k) + ome * (double(ph_p) - ph)) -
                   c2a_m * c.rdnw[k - 1] / (c.c1h[k - 1] * m + c.c2h[k - 1]) * (ope * (r_k - r_km1) + ome * (double(ph) - ph_m)))) +
             dts * kG * msft_inv * (c.rdn[k] * (double(c2a) * alt * t2a_k - double(c2a_m) * alt_m * t2a_km) - c.c1f[k] * mav);
    }
    const int km = nz - 1;
    double wk = w2 + dts * rwt +
                msft_inv * (-H * dts * kG / (c.c1h[km] * m + c.c2h[km]) * (c.rdnw[km] * c.rdnw[km]) * 2.0 * c2a_m * (ope * (r_k - r_km1) + ome * (double(ph) - ph_m)) -
                            dts * kG * (2.0 * c.rdnw[km] * c2a_m * alt_m * t2a_km + c.c1f[k] * mav));
    if (a.q.top_lid) wk = 0;
    return wk;
  };
  // the forward sweep of the tridiagonal solve (KWSolve, small_step.cpp:386), FP64
  auto sweep = [&](int k, double wexp) {
    const idx_t xk = x0 + k * sk;
    const double y = k == 0 ? wexp : (wexp - f.ca[xk] * y_prev) * f.calpha[xk];
    f.ky[xk] = y;
    y_prev = y;
  };
  for (int n = 0; n <= nz; ++n) {
    const idx_t x = x0 + n * sk;
    const bool mk = n < nz;
    Lev l{};
    if (mk) {
      // omega'' at w level n+1 from the running integral through level n (KMuTCol, small_step.cpp:231-247)
      const double dv = divergence(n, x, l);
      if (n + 1 < nz) {
        ww_full = ww_full - c.dnw[n] * (c.c1h[n] * dmdt + dv + c.c1h[n] * mtend) / msfty;
        ww_p1 = ww_full - double(f.ww1[x + sk]);
      } else {
        ww_p1 = 0;
      }
      ts_p1 = n + 1 < nz ? f.t_save[x + sk] : real(0);
      phs_p1 = f.ph_save[x + sk];
      phb_p1 = f.phb(i, j, n + 1);
    }
    // ww, ww_m at w level n
    {
      const real ww1 = f.ww1[x];
      f.ww[x] = ww_0;
      const double sw = (first ? 0.0 : double(f.ww_m[x])) + ww_0;
      put<kHMass>(f.ww_m, g, i, j, n * sk, last ? sw / rn + ww1 : sw);
    }
    double t2a_0 = 0;
    real c2a_0 = 0, alt_0 = 0;
    if (mk) {
      // ru_m, rv_m (KMuTPoint, small_step.cpp:250-262)
      const double su0 = (first ? 0.0 : double(f.ru_m[x])) + l.u0;
      put<kHU>(f.ru_m, g, i, j, n * sk, last ? su0 / rn + (c.c1h[n] * muu0 + c.c2h[n]) * l.us0 / msfuy0 : su0);
      if (!two_d) {
        const double sv0 = (first ? 0.0 : double(f.rv_m[x])) + l.v0;
        put<kHV>(f.rv_m, g, i, j, n * sk, last ? sv0 / rn + (c.c1h[n] * muv0 + c.c2h[n]) * l.vs0 * mvi0 : sv0);
      }
      // theta'' (small_step.cpp:263-276)
      const double ts = ts_0;
      const double wdtn_k = n == 0 ? 0.0 : ww_0 * (c.fnm[n] * ts + c.fnp[n] * ts_m1);
      const double wdtn_kp = n + 1 == nz ? 0.0 : ww_p1 * (c.fnm[n + 1] * ts_p1 + c.fnp[n + 1] * ts);
      double hflux = H * c.rdx * (double(l.u1) * (double(f.t_save[x + 1]) + ts) - double(l.u0) * (double(ts) + f.t_save[x - 1]));
      if (!two_d)
        hflux += H * c.rdy * (double(l.v1) * (double(f.t_save[x + sj]) + ts) - double(l.v0) * (double(ts) + f.t_save[x - sj]));
      const real t2 = f.t2[x];
      const double tn = t2 + msfty * dts * f.t_tend[x];
      const double t2new = tn - dts * msfty * (msftx * hflux + c.rdnw[n] * (wdtn_kp - wdtn_k));
      f.t2[x] = t2new;
      const double tav = H * ((1.0 + epssm) * t2new + (1.0 - epssm) * t2);
      t2a_0 = (tav + c.c1h[n] * muave * kT0) / ((c.c1h[n] * muts + c.c2h[n]) * (kT0 + ts));
      c2a_0 = f.c2a[x];
      alt_0 = f.alt[x];
    }
    // the explicit phi'' R at w level n (KWRhs, small_step.cpp:312-328)
    double R_0 = 0;
    const real ph_p1 = mk ? f.ph2[x + sk] : real(0);
    if (!(n == 0 || (a.q.top_lid && n == nz))) {
      const double wdwn_k = H * (ww_0 + ww_m1) * c.rdnw[n - 1] * (double(phs_0) - phs_m1 + phb_0 - phb_m1);
      double rr = dts * (f.ph_tend[x] + H * kG * (1.0 - epssm) * f.w2[x]);
      if (mk) {
        const double wdwn_kp = H * (ww_p1 + ww_0) * c.rdnw[n] * (double(phs_p1) - phs_0 + phb_p1 - phb_0);
        rr -= dts * (c.fnm[n] * wdwn_kp + c.fnp[n] * wdwn_k);
      }
      R_0 = ph_0 + msfty * rr / (c.c1f[n] * mut + c.c2f[n]);
    }
    f.kr[x] = R_0;
    // the explicit w'' and the forward sweep lag one level (w'' at n-1 needs R at n)
    if (n == 1) sweep(0, w_expl0);
    if (n >= 2) sweep(n - 1, w_explicit(n - 1, R_m1, R_m2, R_0, ph_m1, ph_m2, ph_0, t2a_m1, t2a_m2, c2a_m1, c2a_m2, alt_m1, alt_m2));
    if (n == nz) sweep(nz, w_explicit(nz, R_0, R_m1, 0.0, ph_0, ph_m1, real(0), 0.0, t2a_m1, real(0), c2a_m1, real(0), alt_m1));
    ww_m1 = ww_0;
    ww_0 = ww_p1;
    ts_m1 = ts_0;
    ts_0 = ts_p1;
    phs_m1 = phs_0;
    phs_0 = phs_p1;
    phb_m1 = phb_0;
    phb_0 = phb_p1;
    ph_m2 = ph_m1;
    ph_m1 = ph_0;
    ph_0 = ph_p1;
    R_m2 = R_m1;
    R_m1 = R_0;
    t2a_m2 = t2a_m1;
    t2a_m1 = t2a_0;
    c2a_m2 = c2a_m1;
    c2a_m1 = c2a_0;
    alt_m2 = alt_m1;
    alt_m1 = alt_0;
  }
  // ---------------------------------------------------------------- pass 3: back-substitution, damping, phi'', p'', al''
  const double dampmag = a.q.damp_opt == 3 ? dts * a.q.dampcoef : 0.0;
  double w_next = 0;
  double ph_up = 0;  // phi'' at w level k + 1
  for (int k = nz; k >= 0; --k) {
    const idx_t x = x0 + k * sk;
    const double y = f.ky[x];
    double wk = (k == nz || k == 0) ? y : y - f.cgamma[x] * w_next;
    w_next = wk;  // the back-substitution runs on the undamped values (small_step.cpp:387 before 393-405)
    double phk;
    if (k >= 1) {
      if (dampmag != 0.0) {  // the stage's weight (damp_weight)
        const double dampwt = f.dampwt[x];
        wk = (wk - dampwt * (c.c1f[k] * mut + c.c2f[k]) * f.w_save[x]) / (1.0 + dampwt);
      }
      phk = f.kr[x] + msfty * H * dts * kG * (1.0 + epssm) * wk / (c.c1f[k] * muts + c.c2f[k]);
      put<kHMass>(f.ph2, g, i, j, k * sk, phk);
    } else {
      phk = f.ph2[x];
    }
    f.w2[x] = wk;
    if (k < nz) {  // p'', al'' of mass level k (KWPRho, small_step.cpp:409-414)
      const double m = c.c1h[k] * muts + c.c2h[k];
      const real alt = f.alt[x], ts = f.t_save[x], c2a = f.c2a[x];
      const double al = -((1. / m) * (alt * (c.c1h[k] * mu_new) + c.rdnw[k] * (ph_up - phk)));
      put<kHMass>(f.al, g, i, j, k * sk, al);
      double pp = c2a * (alt * (double(f.t2[x]) - c.c1h[k] * mu_new * ts) / (m * (kT0 + ts)) - al);
      const double ptmp = pp;
      pp = pp + a.q.smdiv * (pp - f.pm1[x]);
      f.pm1[x] = ptmp;
      put<kHMass>(f.p, g, i, j, k * sk, pp);
    }
    ph_up = phk;
  }
}

// ----------------------------------------------------------------------------- batched substep launches
// Every member of a batch in one launch (the acoustic substeps of an ensemble advance together): block b of the
// launch's 1-D grid belongs to the member m with tab[m].start <= b < tab[m + 1].start and is that member's block
// b - tab[m].start; the member's Args and the stage's dts come from device tables the batch keeps.  The blocks of
// one member are contiguous, so the XCD remap keeps a member's neighbouring blocks on one XCD.
__device__ __forceinline__ int batch_member(const BatchMember* tab, int n, int blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].start <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
// advance_uv of every member (point blocks of 64 x 4 over each member's uv range)
MI_POINT_KERNEL k_uv_batch(const Args* as, const BatchMember* tab, int n) {
  const Blk b = xcd_block();
  const int m = batch_member(tab, n, b.x);
  const BatchMember& t = tab[m];
  const int l = b.x - t.start;
  const int bx = l % t.gx, by = (l / t.gx) % t.gy, bz = l / (t.gx * t.gy);
  const int i = t.i0 + bx * kBx + int(threadIdx.x), j = t.j0 + by * int(blockDim.y) + int(threadIdx.y), k = t.k0 + bz;
  if (i >= t.i1 || j >= t.j1) return;
  uv_at<true>(as[m], i, j, k, t.dts);  // the open-boundary forms are taken only where a member's grid is open
}
// the substep's column kernel of every member (blocks of CI columns x KT levels; KT >= nz + 1 of every member)
template <int CI, int KT>
__global__ void __launch_bounds__(CI * KT) k_acoustic_batch(const Args* as, const BatchMember* tab, int n, int it, int ns) {
  extern __shared__ double lds[];
  const Blk b = xcd_block();
  const int m = batch_member(tab, n, b.x);
  const BatchMember& t = tab[m];
  const int l = b.x - t.start;
  ColPos cp;
  cp.ci = int(threadIdx.x) % CI;
  cp.kt = int(threadIdx.x) / CI;
  cp.i = (l % t.gx) * CI + cp.ci;
  cp.j = l / t.gx;
  acoustic_cols<CI, KT>(as[m], cp, lds, t.dts, it, ns);
}

// KSumEdge (small_step.cpp:279-290, 560-569) of an open direction, after each substep's column kernel: the east
// (north) domain face is a real face, its time average summed as the faces of the columns (with specified
// boundaries k_acoustic_ring does it).  ydir 0: faces i = nx of the last column; 1: faces j = ny of the last row.
MI_POINT_KERNEL k_open_sum(Range r, Args a, int it, int ns, int ydir) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const bool first = it == 1, last = it == ns;
  const double rn = double(ns);
  const idx_t sk = g.sk;
  if (ydir == 0) {
    const idx_t y = g.at(i + 1, j, k), y2 = g.at2(i + 1, j);
    const real u1 = f.u2[y], us1 = f.u_save[y];
    const double se = (first ? 0.0 : double(f.ru_m[y])) + u1;
    put<kHU>(f.ru_m, g, i + 1, j, k * sk, last ? se / rn + (c.c1h[k] * f.muu[y2] + c.c2h[k]) * us1 / f.msfuy[y2] : se);
  } else {
    const idx_t y = g.at(i, j + 1, k), y2 = g.at2(i, j + 1);
    const real v1 = f.v2[y], vs1 = f.v_save[y];
    const double sn = (first ? 0.0 : double(f.rv_m[y])) + v1;
    put<kHV>(f.rv_m, g, i, j + 1, k * sk, last ? sn / rn + (c.c1h[k] * f.muv[y2] + c.c2h[k]) * vs1 * f.msfvx_inv[y2] : sn);
  }
}

// ----------------------------------------------------------------------------- the spec zone's substep
// One acoustic substep of the spec zone's columns with specified boundaries (a block per column, a thread per
// level), in the reference's order (SmallStep::run_stage, small_step.cpp:548-593, with LateralBC's updates,
// lateral_bc.cpp:356-387):
//   KMuTCol's ring branch: omega'' left as it stands, summed into ww_m (small_step.cpp:214-220)
//   KMuTPoint's ru_m, rv_m sums (theta'' not advanced there) and KSumEdge's east and north domain faces
//   update_mu_t: theta'', mu'' and muts += dts * tendency
//   spec_bdyupdate_ph (KUpdatePh): phi'' of every level from the stage masses; w waits for zero_grad_bdy
//   KWPRho: p'', al'' with divergence damping
// The interior columns do not read these columns within a substep (every acoustic kernel reads its own
// column, and the spec zone's faces were advanced by k_uv before), so the two kernels are independent.
template <int KT>
__global__ void __launch_bounds__(KT) k_acoustic_ring(Args a, double dts, int it, int ns) {
  __shared__ double PH[KT + 1];
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  int col = int(blockIdx.x), i, j;  // S row, N row, W column, E column
  if (col < nx) {
    i = col;
    j = 0;
  } else if ((col -= nx) < nx) {
    i = col;
    j = ny - 1;
  } else if ((col -= nx) < ny - 2) {
    i = 0;
    j = 1 + col;
  } else {
    i = nx - 1;
    j = 1 + col - (ny - 2);
  }
  const int k = int(threadIdx.x);
  const bool on = k <= nz, mk = k < nz;
  const bool first = it == 1, last = it == ns;
  const double rn = double(ns);
  const idx_t sk = g.sk, sj = g.sj, x2 = g.at2(i, j), x = g.at(i, j, on ? k : nz);
  const double mu_old = f.mu2[x2], mtend = f.mu_tend[x2], muts_old = f.muts[x2];
  double wwcur = 0, wwm = 0, ww1 = 0, ph = 0, phtend = 0, phs = 0;
  double u0 = 0, us0 = 0, v0 = 0, vs0 = 0, rum = 0, rvm = 0, u1 = 0, us1 = 0, v1 = 0, vs1 = 0, rum1 = 0, rvm1 = 0;
  double t2 = 0, ttend = 0, ts = 0, c2a = 0, alt = 0, pm1 = 0;
  if (on) {
    wwcur = f.ww[x];
    if (!first) wwm = f.ww_m[x];
    ww1 = f.ww1[x];
    ph = f.ph2[x];
    phtend = f.ph_tend[x];
    phs = f.ph_save[x];
    if (mk) {
      u0 = f.u2[x];
      us0 = f.u_save[x];
      v0 = f.v2[x];
      vs0 = f.v_save[x];
      t2 = f.t2[x];
      ttend = f.t_tend[x];
      ts = f.t_save[x];
      c2a = f.c2a[x];
      alt = f.alt[x];
      pm1 = f.pm1[x];
      if (i == nx - 1) {
        u1 = f.u2[x + 1];
        us1 = f.u_save[x + 1];
      }
      if (j == ny - 1) {
        v1 = f.v2[x + sj];
        vs1 = f.v_save[x + sj];
      }
      if (!first) {
        rum = f.ru_m[x];
        rvm = f.rv_m[x];
        if (i == nx - 1) rum1 = f.ru_m[x + 1];
        if (j == ny - 1) rvm1 = f.rv_m[x + sj];
      }
    }
  }
  __syncthreads();  // every level has read mu'' and muts before they change
  const double mu_new = mu_old + dts * mtend, muts = muts_old + dts * mtend;
  if (k == 0) {
    put<kHMass>(f.mu2, g, i, j, 0, mu_new);
    put<kHMass>(f.muts, g, i, j, 0, muts);
  }
  double t2new = 0;
  if (on) {
    const double sw = (first ? 0.0 : wwm) + wwcur;
    put<kHMass>(f.ww_m, g, i, j, k * sk, last ? sw / rn + ww1 : sw);
    if (mk) {
      const double muu0 = f.muu[x2], msfuy0 = f.msfuy[x2], muv0 = f.muv[x2], mvi0 = f.msfvx_inv[x2];
      const double su = (first ? 0.0 : rum) + u0;
      put<kHU>(f.ru_m, g, i, j, k * sk, last ? su / rn + (c.c1h[k] * muu0 + c.c2h[k]) * us0 / msfuy0 : su);
      const double sv = (first ? 0.0 : rvm) + v0;
      put<kHV>(f.rv_m, g, i, j, k * sk, last ? sv / rn + (c.c1h[k] * muv0 + c.c2h[k]) * vs0 * mvi0 : sv);
      if (i == nx - 1) {
        const double se = (first ? 0.0 : rum1) + u1;
        put<kHU>(f.ru_m, g, i + 1, j, k * sk,
                 last ? se / rn + (c.c1h[k] * f.muu[x2 + 1] + c.c2h[k]) * us1 / f.msfuy[x2 + 1] : se);
      }
      if (j == ny - 1) {
        const double sn = (first ? 0.0 : rvm1) + v1;
        put<kHV>(f.rv_m, g, i, j + 1, k * sk,
                 last ? sn / rn + (c.c1h[k] * f.muv[x2 + sj] + c.c2h[k]) * vs1 * f.msfvx_inv[x2 + sj] : sn);
      }
      t2new = t2 + dts * ttend;
      f.t2[x] = t2new;
    }
    const double mu_o = muts - dts * mtend;
    const double m_old = c.c1f[k] * mu_o + c.c2f[k], m_new = c.c1f[k] * muts + c.c2f[k];
    const double phk = ph * m_old / m_new + dts * phtend / m_new + phs * (m_old / m_new - 1.0);
    put<kHMass>(f.ph2, g, i, j, k * sk, phk);
    PH[k] = phk;
  }
  __syncthreads();
  if (mk) {
    const double m = c.c1h[k] * muts + c.c2h[k];
    const double al = -((1. / m) * (alt * (c.c1h[k] * mu_new) + c.rdnw[k] * (PH[k + 1] - PH[k])));
    put<kHMass>(f.al, g, i, j, k * sk, al);
    double pp = c2a * (alt * (t2new - c.c1h[k] * mu_new * ts) / (m * (kT0 + ts)) - al);
    const double ptmp = pp;
    pp = pp + a.q.smdiv * (pp - pm1);
    f.pm1[x] = ptmp;
    put<kHMass>(f.p, g, i, j, k * sk, pp);
  }
}

// ----------------------------------------------------------------------------- small_step_finish
// KFinish + KFinishMu (small_step.cpp:419-454) over the (nx+1) x nyv columns and, with prhophi set,
// calc_p_rho_phi (KPRhoPhi, small_step.cpp:456-477) of the mass columns from the recoupled state: the
// reference refreshes halos in between, which KPRhoPhi does not read.
MI_COLUMN_KERNEL k_finish(Args a, int prhophi, double hd_dt) {
  constexpr int kKT = KT;
  extern __shared__ double lds[];
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const ColPos cp = col_pos<CI>();
  const int i = cp.i, j = cp.j, nx = g.nx, ny = g.ny, nz = g.nz, ci = cp.ci;
  const bool on = i <= nx;
  const bool mass = i < nx && j < ny;
  double* PH = lds;
  double* MU = PH + (nz + 1) * CI;
  const idx_t x2 = g.at2(i, j);
  if (on) {
    const double muts = f.muts[x2];
    const double mus_u = j < ny ? H * (muts + f.muts[x2 - 1]) : 0.0;
    const double mus_v = (i < nx && j < g.nyv) ? H * (muts + f.muts[x2 - g.sj]) : 0.0;
    const double m = mass ? f.mut[x2] : 0.0;
    for (int k = cp.kt; k <= nz; k += kKT) {
      const idx_t x = g.at(i, j, k);
      if (j < ny && k < nz)
        put<kHU>(f.u2, g, i, j, k * g.sk, (f.msfuy[x2] * f.u2[x] + f.u_save[x] * (c.c1h[k] * f.muu[x2] + c.c2h[k])) / (c.c1h[k] * mus_u + c.c2h[k]));
      if (i < nx && j < g.nyv && k < nz)
        put<kHV>(f.v2, g, i, j, k * g.sk, (f.msfvx[x2] * f.v2[x] + f.v_save[x] * (c.c1h[k] * f.muv[x2] + c.c2h[k])) / (c.c1h[k] * mus_v + c.c2h[k]));
      if (mass) {
        put<kHMass>(f.w2, g, i, j, k * g.sk, (f.msfty[x2] * f.w2[x] + f.w_save[x] * (c.c1f[k] * m + c.c2f[k])) / (c.c1f[k] * muts + c.c2f[k]));
        const double phk = f.ph2[x] + f.ph_save[x];
        put<kHMass>(f.ph2, g, i, j, k * g.sk, phk);
        LDS(PH, k, ci) = double(real(phk));  // the value as stored (the same as calc_p_rho_phi reads alone)
        put<kHMass>(f.ww, g, i, j, k * g.sk, f.ww[x] + f.ww1[x]);
        if (k < nz) put<kHMass>(f.t2, g, i, j, k * g.sk, (f.t2[x] - (hd_dt != 0.0 ? hd_dt * (c.c1h[k] * m + c.c2h[k]) * f.hdia[x] : 0.0) + f.t_save[x] * (c.c1h[k] * m + c.c2h[k])) / (c.c1h[k] * muts + c.c2h[k]));
      }
    }
    if (cp.kt == 0) {
      if (mass) {
        const double mu = f.mu2[x2] + f.mu_save[x2];
        put<kHMass>(f.mu2, g, i, j, 0, mu);
        MU[ci] = mu;
      }
      if (j < ny) f.muus[x2] = mus_u;
      if (i < nx && j < g.nyv) f.muvs[x2] = mus_v;
    }
  }
  if (!prhophi) return;
  __syncthreads();
  if (on && mass) {
    const double mu2 = MU[ci], muts = f.muts[x2];
    for (int k = cp.kt; k < nz; k += kKT) {
      const idx_t x = g.at(i, j, k);
      const double al = -1.0 / (c.c1h[k] * muts + c.c2h[k]) * (f.alb(i, j, k) * (c.c1h[k] * mu2) + c.rdnw[k] * (LDS(PH, k + 1, ci) - LDS(PH, k, ci)));
      put<kHMass>(f.al, g, i, j, k * g.sk, al);
      const double base = kRd * (kT0 + f.t2[x]) / (kP1000 * (al + f.alb(i, j, k)));
      put<kHMass>(f.p, g, i, j, k * g.sk, kP1000 * mi_libm::pow(base, kCpovcv) - f.pb(i, j, k));
    }
  }
}

// ----------------------------------------------------------------------------- calc_p_rho_phi
// KPRhoPhi (small_step.cpp:456-477), hypsometric_opt = 1, glibc pow
MI_POINT_KERNEL k_p_rho_phi(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double al = -1.0 / (c.c1h[k] * f.muts[x2] + c.c2h[k]) * (f.alb(i, j, k) * (c.c1h[k] * f.mu2[x2]) + c.rdnw[k] * (double(f.ph2[x + g.sk]) - double(f.ph2[x])));
  put<kHMass>(f.al, g, i, j, k * g.sk, al);
  const double base = kRd * (kT0 + f.t2[x]) / (kP1000 * (al + f.alb(i, j, k)));
  put<kHMass>(f.p, g, i, j, k * g.sk, kP1000 * mi_libm::pow(base, kCpovcv) - f.pb(i, j, k));
}

// ----------------------------------------------------------------------------- microphysics after the RK loop
// PhysicsCoupling::step with Kessler warm rain (reference src/dyn/physics_coupling.cpp:262-455,
// src/phys/mp_kessler/kessler.hpp:16-151; WRF moist_physics_prep_em, module_mp_kessler.F,
// moist_physics_finish_em with use_theta_m = 1).  Species 0, 1, 2 are qv, qc, qr.  The libm calls are glibc's
// scalar pow and exp (the reference built with CPPWRF_NO_LIBMVEC), in FP64 in both storage modes.

// moist_physics_prep_em (KPrep, KSaveQc): every mass point
MI_POINT_KERNEL k_mp_prep(Range r, Args a, Scalars s) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  Fields f = a.f;
  const idx_t x = g.at(i, j, k);
  const double zw0 = (double(f.ph2[x]) + f.phb(i, j, k)) / kG;
  const double zw1 = (double(f.ph2[x + g.sk]) + f.phb(i, j, k + 1)) / kG;
  f.dz8w[x] = real(zw1 - zw0);
  f.zmid[x] = real(0.5 * (zw0 + zw1));
  const double q = s.q2[0][x];
  const double th = (double(f.t2[x]) + kT0) / (1.0 + kRvovrd * q);
  f.th_phy[x] = real(th);
  f.hdia[x] = real(th);
  f.qvd[x] = real(q);
  f.mrho[x] = real(1.0 / (double(f.al[x]) + f.alb(i, j, k)));
  f.pii[x] = real(mi_libm::pow((double(f.p[x]) + f.pb(i, j, k)) / kP1000, kRcp));
  f.qcd[x] = s.q2[1][x];
}

// Kessler (KesslerKernel, kessler.hpp:16-151): one thread per column outside the spec zone.  The per-level values
// the reference keeps in column arrays and that do not change (density, the layer spacings, the density ratio of
// the fall speed) are recomputed from the fields with the same expressions; the fall speeds and the falling rain
// stay in thread arrays
__global__ void __launch_bounds__(64) k_kessler(Args a, Scalars s, double dt, int i0, int i1, int j0, int j1) {
  const Blk b = xcd_block();
  const int i = i0 + b.x * 64 + int(threadIdx.x), j = j0 + b.y;
  if (i >= i1 || j >= j1) return;
  const Geo& g = a.g;
  Fields f = a.f;
  real* qv = s.q2[0];
  real* qc = s.q2[1];
  real* qr = s.q2[2];
  const int nz = g.nz;
  const idx_t x0 = g.at(i, j, 0), sk = g.sk;
  double vt[kKesslerMaxLevels], prodk[kKesslerMaxLevels];
  const double xlv = 2.5e6, cp = kCp, ep2 = kRd / 461.6, svp1 = 0.6112, svp2 = 17.67, svp3 = 29.65, svpt0 = 273.15,
               rhowater = 1000.0;
  const double f5 = svp2 * (svpt0 - svp3) * xlv / cp;
  const double c1 = 0.001, c2 = 0.001, c3 = 2.2, c4 = 0.875, max_cr_sedimentation = 0.75;
  auto rho = [&](int k) { return double(f.mrho[x0 + k * sk]); };
  auto rdzk = [&](int k) {
    return k < nz - 1 ? 1.0 / (double(f.zmid[x0 + (k + 1) * sk]) - double(f.zmid[x0 + k * sk]))
                      : 1.0 / (double(f.zmid[x0 + (nz - 1) * sk]) - double(f.zmid[x0 + (nz - 2) * sk]));
  };
  auto rdzw = [&](int k) { return 1.0 / double(f.dz8w[x0 + k * sk]); };
  const double rho0 = rho(0);
  auto vtden = [&](int k) { return __builtin_sqrt(rho0 / rho(k)); };
  auto mx = [](double p, double q) { return p < q ? q : p; };  // std::max(p, q): p on ties and NaN
  auto mn = [](double p, double q) { return q < p ? q : p; };  // std::min(p, q): p on ties and NaN
  double crmax = 0;
  for (int k = 0; k < nz; ++k) {
    prodk[k] = qr[x0 + k * sk];
    const double qrr = mx(0.0, double(qr[x0 + k * sk]) * 0.001 * rho(k));
    vt[k] = 36.34 * mi_libm::pow(qrr, 0.1364) * vtden(k);
    crmax = mx(vt[k] * dt * rdzw(k), crmax);
  }
  int nfall = 1;
  {
    const int n = int(__builtin_round(0.5 + crmax / max_cr_sedimentation));
    nfall = n > 1 ? n : 1;
  }
  double dtfall = dt / double(nfall);
  double time_sediment = dt;
  const idx_t x2 = g.at2(i, j);
  while (nfall > 0) {
    time_sediment = time_sediment - dtfall;
    const double ppt = rho0 * prodk[0] * vt[0] * dtfall / rhowater;
    f.rainncv[x2] = ppt * 1000.0;
    f.rainnc[x2] = f.rainnc[x2] + ppt * 1000.0;
    for (int k = 0; k < nz - 1; ++k) {
      const double factor = dtfall * rdzk(k) / rho(k);
      prodk[k] = prodk[k] - factor * (rho(k) * prodk[k] * vt[k] - rho(k + 1) * prodk[k + 1] * vt[k + 1]);
    }
    {
      const double factor = dtfall * rdzk(nz - 1);
      prodk[nz - 1] = prodk[nz - 1] - factor * prodk[nz - 1] * vt[nz - 1];
    }
    if (nfall > 1) {
      nfall = nfall - 1;
      crmax = 0;
      for (int k = 0; k < nz; ++k) {
        const double qrr = mx(0.0, prodk[k] * 0.001 * rho(k));
        vt[k] = 36.34 * mi_libm::pow(qrr, 0.1364) * vtden(k);
        crmax = mx(vt[k] * time_sediment * rdzw(k), crmax);
      }
      const int n = int(__builtin_round(0.5 + crmax / max_cr_sedimentation));
      const int nfall_new = n > 1 ? n : 1;
      if (nfall_new != nfall) {
        nfall = nfall_new;
        dtfall = time_sediment / double(nfall);
      }
    } else {
      nfall = 0;  // prod = prodk
    }
  }
  for (int k = 0; k < nz; ++k) {
    const idx_t x = x0 + k * sk;
    double qck = qc[x], qrk = qr[x], qvk = qv[x], th = f.th_phy[x];
    const double pk = f.pii[x], rhk = f.mrho[x];
    const double factorn = 1.0 / (1.0 + c3 * dt * mi_libm::pow(mx(0.0, qrk), c4));
    const double qrprod = qck * (1.0 - factorn) + factorn * c1 * dt * mx(qck - c2, 0.0);
    const double rcgs = 0.001 * rhk;
    qck = mx(qck - qrprod, 0.0);
    qrk = (qrk + prodk[k] - qrk);
    qrk = mx(qrk + qrprod, 0.0);
    const double temp = pk * th;
    const double pressure = 1.000e+05 * mi_libm::pow(pk, 1004.0 / 287.0);
    const double gam = 2.5e+06 / (1004.0 * pk);
    const double es = 1000.0 * svp1 * mi_libm::exp(svp2 * (temp - svpt0) / (temp - svp3));
    const double qvs = ep2 * es / (pressure - es);
    const double prod = (qvk - qvs) / (1.0 + pressure / (pressure - es) * qvs * f5 / ((temp - svp3) * (temp - svp3)));
    const double dim_qvs_qv = mx(0.0, qvs - qvk);
    const double ern_num = (1.6 + 124.9 * mi_libm::pow(rcgs * qrk, 0.2046)) * mi_libm::pow(rcgs * qrk, 0.525);
    const double ern_den = 2.55e8 / (pressure * qvs) + 5.4e5;
    const double ern_term1 = dt * (ern_num / ern_den) * (dim_qvs_qv / (rcgs * qvs));
    const double ern_term2 = mx(-prod - qck, 0.0);
    const double ern = mn(ern_term1, mn(ern_term2, qrk));
    const double product = mx(prod, -qck);
    th = th + gam * (product - ern);
    qvk = mx(qvk - product + ern, 0.0);
    qck = qck + product;
    qrk = qrk - ern;
    f.th_phy[x] = real(th);
    qv[x] = real(qvk);
    qc[x] = real(qck);
    qr[x] = real(qrk);
  }
}

// moist_physics_finish_em (KFinish): every mass point; h_diabatic becomes the heating rate the next step's
// dynamics feels
MI_POINT_KERNEL k_mp_finish(Range r, Args a, Scalars s, double dt) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  Fields f = a.f;
  const idx_t x = g.at(i, j, k);
  const double lim = a.q.mp_lim;
  const double th = f.th_phy[x], th_old = f.hdia[x], q = s.q2[0][x], q_old = f.qvd[x];
  double mpten = th - th_old;
  mpten = mpten < lim * dt ? mpten : lim * dt;
  mpten = mpten > -lim * dt ? mpten : -lim * dt;
  const double qvten = q - q_old;
  f.t2[x] = real(th_old * (1.0 + kRvovrd * q_old) + mpten * (1.0 + kRvovrd * q) + kRvovrd * qvten * th - kT0);
  f.hdia[x] = real((mpten * (1.0 + kRvovrd * q) + kRvovrd * qvten * th) / dt);
  f.qvd[x] = real(qvten / dt);
  f.qcd[x] = real((double(s.q2[1][x]) - double(f.qcd[x])) / dt);
}

// ----------------------------------------------------------------------------- sixth-order diffusion
// sixth_order_diffusion (reference src/dyn/diffusion/sixth_order.hpp:13-210; WRF
// module_big_step_utilities_em.F): the flux differences of one direction, then the flux-limited variant
__device__ __forceinline__ double d6(double m3, double m2, double m1, double q0, double p1, double p2) {
  return 10.0 * (q0 - m1) - 5.0 * (p1 - m2) + (p2 - m3);
}
// tendency + tend_x + tend_y of one field at index y (2-D index y2): kind 'u', 'v' or mass ('m' and 'w');
// c1, c2 the level's mass coefficients; mu the column mass (mut for the dynamics, muts for the scalars)
template <char KIND>
__device__ __forceinline__ double sixth_at(double tend, const Args& a, const real* q, const double* mu, idx_t y, idx_t y2,
                                           double c1, double c2, double coef) {
  const Fields& f = a.f;
  const idx_t sj = a.g.sj;
  const int opt = a.q.diff_6th_opt;
  auto lim = [&](double dfl, double diff) { return (opt == 2 && dfl * diff <= 0.0) ? 0.0 : dfl; };
  const double dxp0 = lim(d6(q[y - 3], q[y - 2], q[y - 1], q[y], q[y + 1], q[y + 2]), q[y] - q[y - 1]);
  const double dxp1 = lim(d6(q[y - 2], q[y - 1], q[y], q[y + 1], q[y + 2], q[y + 3]), q[y + 1] - q[y]);
  const double dyp0 = lim(d6(q[y - 3 * sj], q[y - 2 * sj], q[y - sj], q[y], q[y + sj], q[y + 2 * sj]), q[y] - q[y - sj]);
  const double dyp1 = lim(d6(q[y - 2 * sj], q[y - sj], q[y], q[y + sj], q[y + 2 * sj], q[y + 3 * sj]), q[y + sj] - q[y]);
  auto mf = [&](int di, int dj) { return c1 * mu[y2 + di + dj * sj] + c2; };
  double mx0, mx1, my0, my1, sx, sy;
  if (KIND == 'u') {
    mx0 = mf(-1, 0);
    mx1 = mf(0, 0);
    my0 = 0.25 * (mf(-1, -1) + mf(0, -1) + mf(-1, 0) + mf(0, 0));
    my1 = 0.25 * (mf(-1, 0) + mf(0, 0) + mf(-1, 1) + mf(0, 1));
    sx = f.msfux[y2];
    sy = f.msfuy[y2];
  } else if (KIND == 'v') {
    mx0 = 0.25 * (mf(-1, -1) + mf(0, -1) + mf(-1, 0) + mf(0, 0));
    mx1 = 0.25 * (mf(0, -1) + mf(1, -1) + mf(0, 0) + mf(1, 0));
    my0 = mf(0, -1);
    my1 = mf(0, 0);
    sx = f.msfvx[y2];
    sy = f.msfvy[y2];
  } else {
    mx0 = 0.5 * (mf(-1, 0) + mf(0, 0));
    mx1 = 0.5 * (mf(0, 0) + mf(1, 0));
    my0 = 0.5 * (mf(0, -1) + mf(0, 0));
    my1 = 0.5 * (mf(0, 0) + mf(0, 1));
    sx = f.msftx[y2];
    sy = f.msfty[y2];
  }
  const double tx = coef * sx * (mx1 * dxp1 - mx0 * dxp0);
  const double ty = coef * sy * (my1 * dyp1 - my0 * dyp0);
  return tend + tx + ty;
}

// where sixth_order_diffusion reaches (sixth_order.hpp:174-205): everywhere on periodic and symmetric
// sides, three rows in from specified ones ('u' and 'v' two on their own staggered side)
__device__ __forceinline__ bool sixth_in(const Geo& g, int diff, char kind, int i, int j) {
  if (!diff) return false;
  if (g.spec()) {
    const int ie = kind == 'u' ? g.nx - 2 : g.nx - 3, je = kind == 'v' ? g.ny - 2 : g.ny - 3;
    return i >= 3 && i < ie && j >= 3 && j < je;
  }
  // open sides (sixth_order.hpp:179-204): 3 rows in, the end 3 (u, v along their stagger) or 4 rows in; 'v' with
  // open_xs takes the specified limits on all four sides; a periodic x resets the x limits
  const bool ox = g.ox(), oy = g.oy();
  if (!ox && !oy) return true;
  int is = 0, js = 0, ie, je;
  if (kind == 'u') {
    ie = g.nx + 1;
    je = g.ny;
    if (ox) { is = 3; ie = g.nx - 2; }
    if (oy) { js = 3; je = g.ny - 3; }
  } else if (kind == 'v') {
    ie = g.nx;
    je = g.ny + 1;
    if (ox) { is = 3; ie = g.nx - 3; js = 3; je = g.ny - 2; }
    if (oy) { js = 3; je = g.ny - 2; }
  } else {
    ie = g.nx;
    je = g.ny;
    if (ox) { is = 3; ie = g.nx - 3; }
    if (oy) { js = 3; je = g.ny - 3; }
  }
  if (g.bcx == MI_BC_PERIODIC) {
    is = 0;
    ie = kind == 'u' ? g.nx + 1 : g.nx;
  }
  return i >= is && i < ie && j >= js && j < je;
}

// ----------------------------------------------------------------------------- diff_opt 1
// The second-order diffusion along eta of DiffusionDriver with diff_opt 1 (reference
// src/dyn/diffusion/big_step_diffusion.hpp; WRF module_big_step_utilities_em.F): the horizontal eddy viscosities
// xkmh (momenta) and xkhh (theta, scalars) at mass points are averaged onto each face as the reference, term by
// term.  km_opt 1 makes them constants (xkmh = khdif, xkhh = khdif / prandtl) whose sums round as the reference's
// sums of equal terms; km_opt 4 reads the fields the step's first stage sets (k_smag2d).  In an x-z slice the y
// terms are left out: rdy = 0 makes them zeros that leave the sum's bits unchanged (they are added to a tendency
// that is never -0).

// the eddy viscosity at mass point offset (di, dj, dk) from index x: a field (km_opt 4, or damp_opt 1) or the
// constant (km_opt 1)
struct EddyK {
  const real* p;
  double c;
  idx_t x, sj, sk;
  __device__ double operator()(int di, int dj, int dk = 0) const {
    return p ? double(p[x + di + dj * sj + dk * sk]) : c;
  }
};
__device__ __forceinline__ EddyK eddy_k(const Args& a, const real* p, double c, idx_t x) {
  return EddyK{a.q.karr ? p : nullptr, c, x, a.g.sj, a.g.sk};
}

// the loop limits of horizontal_diffusion (big_step_diffusion.hpp:343-414): one row in from open and specified
// sides, a periodic x resetting the x limits
__device__ __forceinline__ bool hdiff_in(const Geo& g, char kind, int i, int j) {
  const bool fx = g.ox() || g.spec(), fy = g.oy() || g.spec();
  int is = fx ? 1 : 0, js = fy ? 1 : 0, ie, je;
  if (kind == 'u') {
    ie = fx ? g.nx : g.nx + 1;
    je = fy ? g.ny - 1 : g.ny;
  } else if (kind == 'v') {
    ie = fx ? g.nx - 1 : g.nx;
    je = fy ? g.ny : g.ny + 1;
  } else {
    ie = fx ? g.nx - 1 : g.nx;
    je = fy ? g.ny - 1 : g.ny;
  }
  if (g.bcx == MI_BC_PERIODIC) {
    is = 0;
    ie = kind == 'u' ? g.nx + 1 : g.nx;
  }
  return i >= is && i < ie && j >= js && j < je;
}

// horizontal_diffusion (KHorizDiffU1 / V1 / W1 / M1, big_step_diffusion.hpp:26-186): tend + the flux differences
// of field q at index x (2-D index x2) with the eddy viscosity K; mu the column mass (mut, or muts for the
// scalars), c1 / c2 the level's mass coefficients (c1f / c2f for 'w')
template <char KIND>
__device__ __forceinline__ double hdiff_at(double tend, const Args& a, const real* q, const double* mu, idx_t x,
                                           idx_t x2, double c1, double c2, const EddyK K) {
  const Fields& f = a.f;
  const idx_t sj = a.g.sj;
  const double rdx = a.c.rdx, rdy = a.c.rdy;
  const bool do_y = !a.g.two_d;
  auto m = [&](int di, int dj) { return c1 * mu[x2 + di + dj * sj] + c2; };
  auto M = [&](const double* p, int di, int dj) { return p[x2 + di + dj * sj]; };
  double mkrdxm, mkrdxp, mrdx, mkrdym = 0.0, mkrdyp = 0.0, mrdy = 0.0;
  if (KIND == 'u') {
    mkrdxm = (M(f.msftx, -1, 0) / M(f.msfty, -1, 0)) * m(-1, 0) * K(-1, 0) * rdx;
    mkrdxp = (M(f.msftx, 0, 0) / M(f.msfty, 0, 0)) * m(0, 0) * K(0, 0) * rdx;
    mrdx = M(f.msfux, 0, 0) * M(f.msfuy, 0, 0) * rdx;
    if (do_y) {
      mkrdym = ((M(f.msfuy, 0, 0) + M(f.msfuy, 0, -1)) / (M(f.msfux, 0, 0) + M(f.msfux, 0, -1))) * 0.25 *
               (m(0, 0) + m(0, -1) + m(-1, -1) + m(-1, 0)) * 0.25 * (K(0, 0) + K(0, -1) + K(-1, -1) + K(-1, 0)) * rdy;
      mkrdyp = ((M(f.msfuy, 0, 0) + M(f.msfuy, 0, 1)) / (M(f.msfux, 0, 0) + M(f.msfux, 0, 1))) * 0.25 *
               (m(0, 0) + m(0, 1) + m(-1, 1) + m(-1, 0)) * 0.25 * (K(0, 0) + K(0, 1) + K(-1, 1) + K(-1, 0)) * rdy;
      mrdy = M(f.msfux, 0, 0) * M(f.msfuy, 0, 0) * rdy;
    }
  } else if (KIND == 'v') {  // no mass factor in y (WRF v4.8.0 lines 2854-2856)
    mkrdxm = ((M(f.msfvx, 0, 0) + M(f.msfvx, -1, 0)) / (M(f.msfvy, 0, 0) + M(f.msfvy, -1, 0))) * 0.25 *
             (m(0, 0) + m(0, -1) + m(-1, -1) + m(-1, 0)) * 0.25 * (K(0, 0) + K(0, -1) + K(-1, -1) + K(-1, 0)) * rdx;
    mkrdxp = ((M(f.msfvx, 0, 0) + M(f.msfvx, 1, 0)) / (M(f.msfvy, 0, 0) + M(f.msfvy, 1, 0))) * 0.25 *
             (m(0, 0) + m(0, -1) + m(1, -1) + m(1, 0)) * 0.25 * (K(0, 0) + K(0, -1) + K(1, -1) + K(1, 0)) * rdx;
    mrdx = M(f.msfvx, 0, 0) * M(f.msfvy, 0, 0) * rdx;
    mkrdym = (M(f.msfty, 0, -1) / M(f.msftx, 0, -1)) * K(0, -1) * rdy;
    mkrdyp = (M(f.msfty, 0, 0) / M(f.msftx, 0, 0)) * K(0, 0) * rdy;
    mrdy = M(f.msfvx, 0, 0) * M(f.msfvy, 0, 0) * rdy;
  } else if (KIND == 'w') {  // K averaged over the mass levels k and k - 1
    mkrdxm = (M(f.msfux, 0, 0) / M(f.msfuy, 0, 0)) * 0.25 * (m(0, 0) + m(-1, 0) + m(0, 0) + m(-1, 0)) * 0.25 *
             (K(0, 0) + K(-1, 0) + K(0, 0, -1) + K(-1, 0, -1)) * rdx;
    mkrdxp = (M(f.msfux, 1, 0) / M(f.msfuy, 1, 0)) * 0.25 * (m(1, 0) + m(0, 0) + m(1, 0) + m(0, 0)) * 0.25 *
             (K(1, 0) + K(0, 0) + K(1, 0, -1) + K(0, 0, -1)) * rdx;
    mrdx = M(f.msftx, 0, 0) * M(f.msfty, 0, 0) * rdx;
    if (do_y) {
      mkrdym = (M(f.msfvy, 0, 0) * M(f.msfvx_inv, 0, 0)) * 0.25 * (m(0, 0) + m(0, -1) + m(0, 0) + m(0, -1)) * 0.25 *
               (K(0, 0) + K(0, -1) + K(0, 0, -1) + K(0, -1, -1)) * rdy;
      mkrdyp = (M(f.msfvy, 0, 1) * M(f.msfvx_inv, 0, 1)) * 0.25 * (m(0, 1) + m(0, 0) + m(0, 1) + m(0, 0)) * 0.25 *
               (K(0, 1) + K(0, 0) + K(0, 1, -1) + K(0, 0, -1)) * rdy;
      mrdy = M(f.msftx, 0, 0) * M(f.msfty, 0, 0) * rdy;
    }
  } else {
    mkrdxm = (M(f.msfux, 0, 0) / M(f.msfuy, 0, 0)) * 0.5 * (K(0, 0) + K(-1, 0)) * 0.5 * (m(0, 0) + m(-1, 0)) * rdx;
    mkrdxp = (M(f.msfux, 1, 0) / M(f.msfuy, 1, 0)) * 0.5 * (K(1, 0) + K(0, 0)) * 0.5 * (m(1, 0) + m(0, 0)) * rdx;
    mrdx = M(f.msftx, 0, 0) * M(f.msfty, 0, 0) * rdx;
    if (do_y) {
      mkrdym = (M(f.msfvy, 0, 0) * M(f.msfvx_inv, 0, 0)) * 0.5 * (K(0, 0) + K(0, -1)) * 0.5 * (m(0, 0) + m(0, -1)) * rdy;
      mkrdyp = (M(f.msfvy, 0, 1) * M(f.msfvx_inv, 0, 1)) * 0.5 * (K(0, 1) + K(0, 0)) * 0.5 * (m(0, 1) + m(0, 0)) * rdy;
      mrdy = M(f.msftx, 0, 0) * M(f.msfty, 0, 0) * rdy;
    }
  }
  const double q0 = q[x];
  const double tx = mrdx * (mkrdxp * (double(q[x + 1]) - q0) - mkrdxm * (q0 - double(q[x - 1])));
  if (!do_y) return tend + tx;
  return tend + (tx + mrdy * (mkrdyp * (double(q[x + sj]) - q0) - mkrdym * (q0 - double(q[x - sj]))));
}

// horizontal_diffusion_3dmp (KHorizDiff3DMP, big_step_diffusion.hpp:188-219): theta minus t_init with xkhh, the
// differences evaluated left to right as WRF
__device__ __forceinline__ double hdiff_3dmp_at(double tend, const Args& a, int i, int j, int k, idx_t x, idx_t x2) {
  const Fields& f = a.f;
  const idx_t sj = a.g.sj;
  const double rdx = a.c.rdx, rdy = a.c.rdy, c1 = a.c.c1h[k], c2 = a.c.c2h[k];
  const EddyK K = eddy_k(a, f.xkhh, a.q.xkhh, x);
  auto m = [&](int di, int dj) { return c1 * f.mut[x2 + di + dj * sj] + c2; };
  auto M = [&](const double* p, int di, int dj) { return p[x2 + di + dj * sj]; };
  auto Q = [&](int di, int dj) { return double(f.t2[x + di + dj * sj]); };
  auto B = [&](int di, int dj) { return f.t_init(i + di, j + dj, k); };
  const double mkrdxm = (M(f.msfux, 0, 0) / M(f.msfuy, 0, 0)) * 0.5 * (K(0, 0) + K(-1, 0)) * 0.5 * (m(0, 0) + m(-1, 0)) * rdx;
  const double mkrdxp = (M(f.msfux, 1, 0) / M(f.msfuy, 1, 0)) * 0.5 * (K(1, 0) + K(0, 0)) * 0.5 * (m(1, 0) + m(0, 0)) * rdx;
  const double mrdx = M(f.msftx, 0, 0) * M(f.msfty, 0, 0) * rdx;
  const double dxp = Q(1, 0) - Q(0, 0) - B(1, 0) + B(0, 0);
  const double dxm = Q(0, 0) - Q(-1, 0) - B(0, 0) + B(-1, 0);
  const double tx = mrdx * (mkrdxp * dxp - mkrdxm * dxm);
  if (a.g.two_d) return tend + tx;
  const double mkrdym =
      (M(f.msfvy, 0, 0) * M(f.msfvx_inv, 0, 0)) * 0.5 * (K(0, 0) + K(0, -1)) * 0.5 * (m(0, 0) + m(0, -1)) * rdy;
  const double mkrdyp =
      (M(f.msfvy, 0, 1) * M(f.msfvx_inv, 0, 1)) * 0.5 * (K(0, 1) + K(0, 0)) * 0.5 * (m(0, 1) + m(0, 0)) * rdy;
  const double mrdy = M(f.msftx, 0, 0) * M(f.msfty, 0, 0) * rdy;
  const double dyp = Q(0, 1) - Q(0, 0) - B(0, 1) + B(0, 0);
  const double dym = Q(0, 0) - Q(0, -1) - B(0, 0) + B(0, -1);
  return tend + (tx + mrdy * (mkrdyp * dyp - mkrdym * dym));
}

// vertical_diffusion_u / _v (KVertDiffU1 / V1, big_step_diffusion.hpp:283-338): at a face of level k; vflux[kk] (kk
// < nz - 1) the flux between levels kk and kk + 1 with alt averaged onto the face and the base profile subtracted;
// no tendency at the top level
template <char KIND>
__device__ __forceinline__ double vdiff_face(double tend, const Args& a, idx_t x, idx_t x2, int k) {
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = a.g.nz;
  const idx_t sk = a.g.sk, o = KIND == 'u' ? 1 : a.g.sj;  // o: the mass point across the face
  const real* q = KIND == 'u' ? f.u2 : f.v2;
  const double* bp = KIND == 'u' ? f.u_base : f.v_base;
  const double* mu = KIND == 'u' ? f.muu : f.muv;
  const double kv = a.q.kvdif;
  auto AL = [&](idx_t y) { return double(f.alt[y]); };
  auto vf = [&](int kk) {
    const idx_t y = x + idx_t(kk - k) * sk;
    const double av = 0.25 * (AL(y) + AL(y - o) + AL(y + sk) + AL(y - o + sk));
    const double dp = double(q[y + sk]) - double(q[y]) - bp[kk + 1] + bp[kk];
    return kv * c.rdn[kk + 1] / av * dp;
  };
  if (k >= nz - 1) return tend;
  const double mm = c.c1h[k] * mu[x2] + c.c2h[k];
  const double av = 0.5 * (AL(x - o) + AL(x));
  if (k == 0) {
    const double v0 = vf(0);
    return tend + kG * kG * c.rdnw[0] / mm / av * (v0 - v0);
  }
  return tend + kG * kG * c.rdnw[k] / mm / av * (vf(k) - vf(k - 1));
}

// vertical_diffusion 'w' (KVertDiffW1, big_step_diffusion.hpp:225-245): w level k in 1..nz-1, the fluxes at the
// mass levels (0 at the top one)
__device__ __forceinline__ double vdiff_w(double tend, const Args& a, idx_t x, idx_t x2, int k) {
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = a.g.nz;
  const idx_t sk = a.g.sk;
  const double kv = a.q.kvdif;
  auto vf = [&](int kk) {
    if (kk >= nz - 1) return 0.0;
    const idx_t y = x + idx_t(kk - k) * sk;
    return (kv / double(f.alt[y])) * c.rdnw[kk] * (double(f.w2[y + sk]) - double(f.w2[y]));
  };
  const double mm = c.c1f[k] * f.mut[x2] + c.c2f[k];
  const double av = 0.5 * (double(f.alt[x]) + double(f.alt[x - sk]));
  return tend + c.rdn[k] * kG * kG / mm / av * (vf(k) - vf(k - 1));
}

// vertical_diffusion 'm' and vertical_diffusion_mp (KVertDiffM1 / MP1, big_step_diffusion.hpp:247-281): mass level
// k of field q; base(kk) the base state subtracted from the differences (mp) when BASE
template <bool BASE, class B>
__device__ __forceinline__ double vdiff_m(double tend, const Args& a, const real* q, const double* mu, B base, double kv,
                                          idx_t x, idx_t x2, int k) {
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = a.g.nz;
  const idx_t sk = a.g.sk;
  auto vf = [&](int kk) {
    if (kk >= nz - 1) return 0.0;
    const idx_t y = x + idx_t(kk - k) * sk;
    const double dq = BASE ? double(q[y + sk]) - double(q[y]) - base(kk + 1) + base(kk) : double(q[y + sk]) - double(q[y]);
    return kv * c.rdn[kk + 1] / (0.5 * (double(f.alt[y]) + double(f.alt[y + sk]))) * dq;
  };
  const double mm = c.c1h[k] * mu[x2] + c.c2h[k];
  const double al = f.alt[x];
  if (k == 0) {
    const double v0 = vf(0);
    return tend + kG * kG / mm / al * c.rdnw[0] * (v0 - v0);
  }
  return tend + kG * kG / mm / al * c.rdnw[k] * (vf(k) - vf(k - 1));
}

// ----------------------------------------------------------------------------- km_opt 4 (2-D Smagorinsky)
// DiffusionDriver::prepare with diff_opt 1, km_opt 4 at the first RK stage (reference diffusion_driver.cpp:92-138):
// the coordinate slopes zx, zy and rdzw of compute_diff_metrics (diff_metrics.hpp:20-138, with the end faces of
// set_physical_bc3d), the horizontal deformations of cal_deform_and_div (deform_div.hpp:20-170, 234-311) and the
// eddy viscosities of KSmag2dKm (km_kh.hpp:47-102, 586-629).  The metrics are evaluated where they are read, from
// ph and phb with the reference's expressions; defor12 is stored at the vorticity points by one launch, defor11
// and defor22 are evaluated by the launch that sets xkmh and xkhh from it.  Every value is the reference's, operand
// for operand, in FP64 arithmetic.

// zx at u point (i, j), w level k in 0..nz: 0 on the boundary faces of a non-periodic x, except the end face of a
// specified or open one (the last inner face copied); a periodic x leaves the top level 0 on its boundary faces
__device__ __forceinline__ double diff_zx(const Args& a, int i, int j, int k) {
  const Geo& g = a.g;
  const int nx = g.nx;
  const bool px = g.bcx == MI_BC_PERIODIC, cx = g.spec() || g.bcx == MI_BC_OPEN;
  if (px && k == g.nz && (i == 0 || i == nx)) return 0.0;
  if (!px && (i == 0 || (i == nx && !cx))) return 0.0;
  const int i0 = (i == nx) ? (px ? 0 : nx - 1) : i, im1 = (i == 0) ? nx - 1 : (i == nx && !px ? nx - 2 : i - 1);
  double v = a.c.rdx * (a.f.phb(i0, j, k) - a.f.phb(im1, j, k)) / kG;
  v += a.c.rdx * (double(a.f.ph2[g.at(i0, j, k)]) - double(a.f.ph2[g.at(im1, j, k)])) / kG;
  return v;
}
// zy at v point (i, j), w level k: the same rules along y (an x-z slice has no open y and no y differences)
__device__ __forceinline__ double diff_zy(const Args& a, int i, int j, int k) {
  const Geo& g = a.g;
  const int ny = g.ny;
  const bool py = g.bcy == MI_BC_PERIODIC, cy = g.spec() || g.oy();
  if (py && k == g.nz && (j == 0 || j == ny)) return 0.0;
  if (!py && (j == 0 || (j == ny && !cy))) return 0.0;
  const int j0 = (j == ny) ? (py ? 0 : ny - 1) : j, jm1 = (j == 0) ? ny - 1 : (j == ny && !py ? ny - 2 : j - 1);
  double v = a.c.rdy * (a.f.phb(i, j0, k) - a.f.phb(i, jm1, k)) / kG;
  v += a.c.rdy * (double(a.f.ph2[g.at(i, j0, k)]) - double(a.f.ph2[g.at(i, jm1, k)])) / kG;
  return v;
}
// rdzw at mass point (i, j, k): 1 / the layer's thickness from the full geopotential
__device__ __forceinline__ double diff_rdzw(const Args& a, int i, int j, int k) {
  const idx_t x = a.g.at(i, j, k);
  const double z0 = (double(a.f.ph2[x]) + a.f.phb(i, j, k)) / kG;
  const double z1 = (double(a.f.ph2[x + a.g.sk]) + a.f.phb(i, j, k + 1)) / kG;
  return 1.0 / (z1 - z0);
}
// the average onto w level k (0..nz) of two columns' values h0(kk), h1(kk) at the mass levels (the hatavg of
// KDeformDiagAndDiv and KDeform12): the surface and top levels extrapolated
template <class H0, class H1>
__device__ __forceinline__ double hat_avg(const Lv& c, int nz, int k, const H0& h0, const H1& h1) {
  if (k == 0) return 0.5 * (c.cf1 * h0(0) + c.cf2 * h0(1) + c.cf3 * h0(2) + c.cf1 * h1(0) + c.cf2 * h1(1) + c.cf3 * h1(2));
  if (k == nz) {
    const double cft2 = -0.5 * c.dnw[nz - 1] / c.dn[nz - 1], cft1 = 1.0 - cft2;
    return 0.5 * (cft1 * (h0(nz - 1) + h1(nz - 1)) + cft2 * (h0(nz - 2) + h1(nz - 2)));
  }
  return 0.5 * (c.fnm[k] * (h0(k) + h1(k)) + c.fnp[k] * (h0(k - 1) + h1(k - 1)));
}

// KDeform12 at vorticity point (i, j), level k (i in 0..nx, j in 0..ny; one row in an x-z slice).  A periodic
// direction takes the start row's values at its end row (the reference computes and then overwrites them by that
// copy); on open and specified sides the edge rows are the reference's copies, which KSmag2dKm never reads, and
// are left as computed here.
MI_POINT_KERNEL k_deform12(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const bool px = g.bcx == MI_BC_PERIODIC, py = g.bcy == MI_BC_PERIODIC;
  const int ii = (px && i == nx) ? 0 : i, jj = (py && j == ny) ? 0 : j;
  const int im1 = (ii == 0) ? (px ? nx - 1 : 0) : ii - 1, i0 = (ii == nx) ? nx - 1 : ii;
  const int jm1 = (jj == 0) ? (py ? ny - 1 : 0) : jj - 1, j0 = (jj == ny) ? ny - 1 : jj;
  const idx_t sk = g.sk;
  const double mm = 0.25 * (f.msfux[g.at2(ii, jm1)] + f.msfux[g.at2(ii, j0)]) *
                    (f.msfvy[g.at2(im1, jj)] + f.msfvy[g.at2(i0, jj)]);
  const idx_t xum = g.at(ii, jm1, 0), xup = g.at(ii, j0, 0), xvm = g.at(im1, jj, 0), xvp = g.at(i0, jj, 0);
  const double fum = f.msfux[g.at2(ii, jm1)], fup = f.msfux[g.at2(ii, j0)];
  const double fvm = f.msfvy[g.at2(im1, jj)], fvp = f.msfvy[g.at2(i0, jj)];
  auto hum = [&](int kk) { return double(f.u2[xum + kk * sk]) / fum; };
  auto hup = [&](int kk) { return double(f.u2[xup + kk * sk]) / fup; };
  auto hvm = [&](int kk) { return double(f.v2[xvm + kk * sk]) / fvm; };
  auto hvp = [&](int kk) { return double(f.v2[xvp + kk * sk]) / fvp; };
  const double tzy = 0.25 * (diff_zy(a, im1, jj, k) + diff_zy(a, i0, jj, k) + diff_zy(a, im1, jj, k + 1) +
                             diff_zy(a, i0, jj, k + 1));
  const double tu = (hat_avg(c, nz, k + 1, hum, hup) - hat_avg(c, nz, k, hum, hup)) * 0.25 * tzy *
                    (diff_rdzw(a, i0, j0, k) + diff_rdzw(a, im1, j0, k) + diff_rdzw(a, im1, jm1, k) +
                     diff_rdzw(a, i0, jm1, k));
  const double term_u = mm * (c.rdy * (hup(k) - hum(k)) - tu);
  const double tzx = 0.25 * (diff_zx(a, ii, jm1, k) + diff_zx(a, ii, j0, k) + diff_zx(a, ii, jm1, k + 1) +
                             diff_zx(a, ii, j0, k + 1));
  const double tv = (hat_avg(c, nz, k + 1, hvm, hvp) - hat_avg(c, nz, k, hvm, hvp)) * 0.25 * tzx *
                    (diff_rdzw(a, i0, j0, k) + diff_rdzw(a, i0, jm1, k) + diff_rdzw(a, im1, jm1, k) +
                     diff_rdzw(a, im1, j0, k));
  const double term_v = mm * (c.rdx * (hvp(k) - hvm(k)) - tv);
  f.d12[g.at(i, j, k)] = term_u + term_v;
}

// KDeformDiagAndDiv's defor11, defor22 at mass point (i, j, k), then KSmag2dKm with diff_opt 1: xkmh = the
// Smagorinsky K (capped at 10 times the grid length), xkhh = xkmh / prandtl.  Launched one row in from open and
// specified sides (a periodic x over every column), as calculate_km_kh; the rows outside keep 0.
MI_POINT_KERNEL k_smag2d(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j), sj = g.sj, sk = g.sk;
  const double mm = f.msftx[x2] * f.msfty[x2];
  const double fu0 = f.msfuy[x2], fu1 = f.msfuy[x2 + 1], fv0 = f.msfvx[x2], fv1 = f.msfvx[x2 + sj];
  const idx_t x0 = x - idx_t(k) * sk;
  auto hu0 = [&](int kk) { return double(f.u2[x0 + kk * sk]) / fu0; };
  auto hu1 = [&](int kk) { return double(f.u2[x0 + 1 + kk * sk]) / fu1; };
  auto hv0 = [&](int kk) { return double(f.v2[x0 + kk * sk]) / fv0; };
  auto hv1 = [&](int kk) { return double(f.v2[x0 + sj + kk * sk]) / fv1; };
  const double rz = diff_rdzw(a, i, j, k);
  const double tzx = 0.25 * (diff_zx(a, i, j, k) + diff_zx(a, i + 1, j, k) + diff_zx(a, i, j, k + 1) +
                             diff_zx(a, i + 1, j, k + 1));
  const double dudx = mm * (c.rdx * (hu1(k) - hu0(k)) - (hat_avg(c, nz, k + 1, hu0, hu1) - hat_avg(c, nz, k, hu0, hu1)) * tzx * rz);
  const double d11 = 2.0 * dudx;
  const int jp = g.two_d ? j : j + 1;  // an x-z slice reads its one row (zero j stride)
  const double tzy = 0.25 * (diff_zy(a, i, j, k) + diff_zy(a, i, jp, k) + diff_zy(a, i, j, k + 1) +
                             diff_zy(a, i, jp, k + 1));
  const double dvdy = mm * (c.rdy * (hv1(k) - hv0(k)) - (hat_avg(c, nz, k + 1, hv0, hv1) - hat_avg(c, nz, k, hv0, hv1)) * tzy * rz);
  const double d22 = 2.0 * dvdy;
  const real* D = f.d12;
  const double dd = d11 - d22;
  double def2 = 0.25 * (dd * dd);
  const double d12a = 0.25 * (double(D[x]) + double(D[x + sj]) + double(D[x + 1]) + double(D[x + 1 + sj]));
  def2 += d12a * d12a;
  const double mlen = __builtin_sqrt(a.q.dx / f.msftx[x2] * a.q.dy / f.msfty[x2]);
  const double tmp = __builtin_sqrt(def2);
  double kmh = a.q.c_s * a.q.c_s * mlen * mlen * tmp;
  kmh = kmh < 10.0 * mlen ? kmh : 10.0 * mlen;
  f.xkmh[x] = kmh;
  f.xkhh[x] = kmh / a.q.prandtl;
}

// rdz at mass point (i, j, k): 2 / the distance between the w levels k - 1 and k + 1 (the first layer's thickness
// at k = 0), from the full geopotential
__device__ __forceinline__ double diff_rdz(const Args& a, int i, int j, int k) {
  const idx_t x = a.g.at(i, j, k), sk = a.g.sk;
  const int km = k == 0 ? 0 : k - 1;
  const double z1 = (double(a.f.ph2[x + sk]) + a.f.phb(i, j, k + 1)) / kG;
  const double z0 = (double(a.f.ph2[x + (km - k) * sk]) + a.f.phb(i, j, km)) / kG;
  return 2.0 / (z1 - z0);
}

// ----------------------------------------------------------------------------- damp_opt 1
// cal_dampkm (reference km_kh.hpp:485-568) after the eddy viscosities of the step: in each column (one row in from
// the relaxation zone of specified sides) xkmh and xkhh are raised to kmmax cos^2(pi/2 min(1, depth / zdamp))
// dampcoef, kmmax = min(dx, dy)^2 / dt of the column, the depth below the top summed down from half the top
// layer in the reference's order.  With km_opt 1 the launch also writes the constants (KIsotropicKm) at every mass
// point.  The vertical K it also raises is not read by diff_opt 1.
MI_POINT_KERNEL k_dampkm(Range r, Args a, int is, int ie, int js, int je) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  double kh = a.q.km4 ? double(f.xkmh[x]) : a.q.xkmh;
  double kt = a.q.km4 ? double(f.xkhh[x]) : a.q.xkhh;
  if (i >= is && i < ie && j >= js && j < je) {
    const int nz = g.nz;
    double deltaz = 0.5 * (1.0 / diff_rdzw(a, i, j, nz - 1));
    for (int kk = nz - 2; kk >= k; --kk) deltaz = deltaz + 1.0 / diff_rdz(a, i, j, kk);
    const double dx_m = a.q.dx / f.msftx[x2], dy_m = a.q.dy / f.msfty[x2];
    const double ds = dx_m < dy_m ? dx_m : dy_m;
    const double kmmax = ds * ds / a.q.dt;
    const double ratio = deltaz / a.q.zdamp;
    const double tmp = ratio < 1.0 ? ratio : 1.0;
    constexpr double piconst = 3.1415926535897932384626433, degrad = piconst / 180.0, degrad90 = degrad * 90.0;
    const double c = mi_libm::cos(degrad90 * tmp);
    const double dampk = c * c * kmmax * a.q.dampcoef;
    if (dampk > kh) kh = dampk;
    if (dampk > kt) kt = dampk;
  }
  f.xkmh[x] = kh;
  f.xkhh[x] = kt;
}

// ----------------------------------------------------------------------------- damp_opt 2
// rk_rayleigh_damp (reference rayleigh_damp.hpp:14-259) into the held tendencies after the diffusion: in the top
// zdamp metres u, v, theta relax towards their base profiles interpolated to each point's height and w towards 0,
// with the coefficient sin^2(pi/2 (1 - min(1, depth / zdamp))) dampcoef.  The damped levels of a u, v or theta
// column run from the top down to the one above the first level below the layer (level 1 if none is below): a
// thread scans its column from the top down to its own level, as the reference's loop does.
// true when mass level k of a column is damped; zk its height zl(k) then
template <class ZL>
__device__ __forceinline__ bool rayleigh_damped(int k, int ktf, double ztop, double zd, const ZL& zl, double& zk) {
  for (int kk = ktf; kk >= k; --kk) {
    const double z = zl(kk);
    if (!(z >= (ztop - zd))) return false;
    zk = z;
  }
  return k >= 1;
}
// the base profile q_base interpolated (or extrapolated at the ends) to height z over z_base
__device__ __forceinline__ double rayleigh_base(const double* qb, const double* zb, int ktf, double z) {
  int k2 = ktf;
  while (k2 > 0 && zb[k2] > z) k2 = k2 - 1;
  if (k2 + 1 > ktf) return qb[k2] + (qb[k2] - qb[k2 - 1]) * (z - zb[k2]) / (zb[k2] - zb[k2 - 1]);
  return qb[k2] + (qb[k2 + 1] - qb[k2]) * (z - zb[k2]) / (zb[k2 + 1] - zb[k2]);
}
MI_POINT_KERNEL k_rayleigh(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nx = g.nx, ny = g.ny, nz = g.nz, ktf = nz - 1;
  const bool px = g.bcx == MI_BC_PERIODIC, py = g.bcy == MI_BC_PERIODIC;
  const double zd = a.q.zdamp, dc = a.q.dampcoef, pii = 3.141592653589793;  // 2 asin(1)
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  auto PH = [&](int ii, int jj, int kk) { return double(f.ph2[g.at(ii, jj, kk)]); };
  auto coef = [&](double ztop, double z) {
    const double ratio = (ztop - z) / zd;
    const double m = ratio < 1.0 ? ratio : 1.0;
    const double s = mi_libm::sin(0.5 * pii * (1.0 - m));
    return s * s;
  };
  if (j < ny && k < nz) {  // KRayleighDampU: u faces i = 0..nx
    const int i0 = (i == nx) ? (px ? 0 : nx - 1) : i, im1 = (i == 0) ? (px ? nx - 1 : 0) : (i - 1);
    const double ztop = 0.5 * (f.phb(i0, j, nz) + f.phb(im1, j, nz) + PH(i0, j, nz) + PH(im1, j, nz)) / kG;
    auto zl = [&](int kk) {
      return 0.25 * (f.phb(i0, j, kk) + f.phb(i0, j, kk + 1) + f.phb(im1, j, kk) + f.phb(im1, j, kk + 1) +
                     PH(i0, j, kk) + PH(i0, j, kk + 1) + PH(im1, j, kk) + PH(im1, j, kk + 1)) / kG;
    };
    double z00 = 0;
    if (rayleigh_damped(k, ktf, ztop, zd, zl, z00)) {
      const double u00 = rayleigh_base(f.u_base, f.z_base, ktf, z00);
      f.ru_tendf[x] = double(f.ru_tendf[x]) - (c.c1h[k] * f.muu[x2] + c.c2h[k]) * (coef(ztop, z00) * dc) *
                                                  (double(f.u2[x]) - u00);
    }
  }
  if (!g.two_d && i < nx && j <= ny && k < nz) {  // KRayleighDampV: v faces j = 0..ny (an x-z slice holds no v)
    const int j0 = (j == ny) ? (py ? 0 : ny - 1) : j, jm1 = (j == 0) ? (py ? ny - 1 : 0) : (j - 1);
    const double ztop = 0.5 * (f.phb(i, j0, nz) + f.phb(i, jm1, nz) + PH(i, j0, nz) + PH(i, jm1, nz)) / kG;
    auto zl = [&](int kk) {
      return 0.25 * (f.phb(i, j0, kk) + f.phb(i, j0, kk + 1) + f.phb(i, jm1, kk) + f.phb(i, jm1, kk + 1) +
                     PH(i, j0, kk) + PH(i, j0, kk + 1) + PH(i, jm1, kk) + PH(i, jm1, kk + 1)) / kG;
    };
    double z00 = 0;
    if (rayleigh_damped(k, ktf, ztop, zd, zl, z00)) {
      const double v00 = rayleigh_base(f.v_base, f.z_base, ktf, z00);
      f.rv_tendf[x] = double(f.rv_tendf[x]) - (c.c1h[k] * f.muv[x2] + c.c2h[k]) * (coef(ztop, z00) * dc) *
                                                  (double(f.v2[x]) - v00);
    }
  }
  if (i < nx && j < ny) {  // KRayleighDampWT
    const double ztop = (f.phb(i, j, nz) + PH(i, j, nz)) / kG;
    const double zw = (f.phb(i, j, k) + PH(i, j, k)) / kG;
    if (zw >= (ztop - zd)) {  // w at every level of the layer, the scalar sine
      const double s2 = coef(ztop, zw);
      f.rw_tendf[x] = double(f.rw_tendf[x]) - (c.c1f[k] * f.mut[x2] + c.c2f[k]) * (s2 * dc) * double(f.w2[x]);
    }
    if (k < nz) {
      auto zl = [&](int kk) { return 0.5 * (f.phb(i, j, kk) + f.phb(i, j, kk + 1) + PH(i, j, kk) + PH(i, j, kk + 1)) / kG; };
      double z00 = 0;
      if (rayleigh_damped(k, ktf, ztop, zd, zl, z00)) {
        const double t00 = rayleigh_base(f.t_base, f.z_base, ktf, z00);
        f.t_tendf[x] = double(f.t_tendf[x]) - (c.c1h[k] * f.mut[x2] + c.c2h[k]) * (coef(ztop, z00) * dc) *
                                                 (double(f.t2[x]) - t00);
      }
    }
  }
}

// zero_tendf + DiffusionDriver::dynamics (diffusion_driver.cpp:266-281): the held tendencies of u, v, w,
// theta (and ph with specified boundaries) over the (nx+1) x nyv x (nz+1) points, 0 where
// sixth_order_diffusion does not reach (w levels 0, nz; the specified edges; everywhere without it)
MI_POINT_KERNEL k_sixth_dyn(Range r, Args a, double coef, int mode) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nx = g.nx, ny = g.ny, nz = g.nz, diff = mode != kSixthZeroOnly && a.q.diff_6th_opt != 0;
  const bool d2 = mode != kSixthZeroOnly && a.q.diff_opt == 1;
  const bool keep = mode == kSixthKeep;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  // tendency(i,j,k) = tendency + the diff_opt 1 terms (horizontal, then vertical) + tend_x + tend_y of the
  // sixth-order diffusion, with the tendency just zeroed, or holding 0 + the physics tendency of part2
  // (update_phy_ten adds into the zeroed field)
  if (j < ny && k < nz) {
    double t = keep ? 0.0 + double(f.ru_tendf[x]) : 0.0;
    if (d2) {
      if (hdiff_in(g, 'u', i, j)) t = hdiff_at<'u'>(t, a, f.u2, f.mut, x, x2, c.c1h[k], c.c2h[k], eddy_k(a, f.xkmh, a.q.xkmh, x));
      // vertical_diffusion_u: all j, i one in from open / specified sides (a periodic x: all faces)
      const bool fx = (g.ox() || g.spec()) && g.bcx != MI_BC_PERIODIC;
      if (i >= (fx ? 1 : 0) && i < (fx ? nx : nx + 1)) t = vdiff_face<'u'>(t, a, x, x2, k);
    }
    f.ru_tendf[x] = sixth_in(g, diff, 'u', i, j) ? sixth_at<'u'>(t, a, f.u2, f.mut, x, x2, c.c1h[k], c.c2h[k], coef) : t;
  }
  if (!g.two_d && i < nx && j < g.nyv && k < nz) {
    double t = keep ? 0.0 + double(f.rv_tendf[x]) : 0.0;
    if (d2) {
      if (hdiff_in(g, 'v', i, j)) t = hdiff_at<'v'>(t, a, f.v2, f.mut, x, x2, c.c1h[k], c.c2h[k], eddy_k(a, f.xkmh, a.q.xkmh, x));
      const bool fy = g.oy() || g.spec();  // vertical_diffusion_v: all i, j one in from open / specified sides
      if (j >= (fy ? 1 : 0) && j < (fy ? ny : ny + 1)) t = vdiff_face<'v'>(t, a, x, x2, k);
    }
    f.rv_tendf[x] = sixth_in(g, diff, 'v', i, j) ? sixth_at<'v'>(t, a, f.v2, f.mut, x, x2, c.c1h[k], c.c2h[k], coef) : t;
  }
  if (i < nx && j < ny) {
    const bool in = sixth_in(g, diff, 'm', i, j), hin = d2 && hdiff_in(g, 'm', i, j);
    double w = keep ? 0.0 + double(f.rw_tendf[x]) : 0.0;
    if (d2 && k >= 1 && k < nz) {
      if (hin) w = hdiff_at<'w'>(w, a, f.w2, f.mut, x, x2, c.c1f[k], c.c2f[k], eddy_k(a, f.xkmh, a.q.xkmh, x));
      w = vdiff_w(w, a, x, x2, k);
    }
    f.rw_tendf[x] = (in && k >= 1 && k < nz) ? sixth_at<'m'>(w, a, f.w2, f.mut, x, x2, c.c1f[k], c.c2f[k], coef) : w;
    if (k < nz) {
      double t = keep ? 0.0 + double(f.t_tendf[x]) : 0.0;
      if (d2) {  // horizontal_diffusion_3dmp, then vertical_diffusion_mp with kvdq = 3 kvdif (WRF rk_tendency)
        if (hin) t = hdiff_3dmp_at(t, a, i, j, k, x, x2);
        const BField tb = f.t_init;
        t = vdiff_m<true>(t, a, f.t2, f.mut, [&](int kk) { return tb(i, j, kk); }, 3.0 * a.q.kvdif, x, x2, k);
      }
      f.t_tendf[x] = in ? sixth_at<'m'>(t, a, f.t2, f.mut, x, x2, c.c1h[k], c.c2h[k], coef) : t;
    }
    if (a.q.spec && !keep) f.ph_tendf[x] = 0.0;
  }
}

// zero_tendf, a column physics plugin's tendency, then DiffusionDriver::scalar (diffusion_driver.cpp:297-327;
// mut there is muts, dt the first stage's).  Without a plugin zero_tendf leaves q_tendf at 0 (large_step.cpp:814,
// 888-893): what the caller holds in it is not read.
MI_POINT_KERNEL k_sixth_sc(Range r, Args a, Scalars s, int n_sc, double coef) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Lv& c = a.c;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const bool in = sixth_in(g, a.q.diff_6th_opt != 0, 'm', i, j);
  const bool d2 = a.q.diff_opt == 1, hin = d2 && hdiff_in(g, 'm', i, j);
  const double* qvb = a.f.qv_base;
#pragma unroll 1
  for (int n = 0; n < n_sc; ++n) {
    double t = 0.0 + (a.q.qtf_on ? double(s.qtf[n][x]) : 0.0);
    if (d2) {  // horizontal_diffusion 'm' with xkhh, then vertical_diffusion(_mp for qv) with kvdq; mut is muts
      if (hin) t = hdiff_at<'m'>(t, a, s.q2[n], a.f.muts, x, x2, c.c1h[k], c.c2h[k], eddy_k(a, a.f.xkhh, a.q.xkhh, x));
      if (n == 0 && s.n_moist > 0)
        t = vdiff_m<true>(t, a, s.q2[n], a.f.muts, [&](int kk) { return qvb[kk]; }, a.q.kvdq, x, x2, k);
      else
        t = vdiff_m<false>(t, a, s.q2[n], a.f.muts, [](int) { return 0.0; }, a.q.kvdq, x, x2, k);
    }
    s.qtw[n][x] = in ? sixth_at<'m'>(t, a, s.q2[n], a.f.muts, x, x2, c.c1h[k], c.c2h[k], coef) : t;
  }
}

// ----------------------------------------------------------------------------- set_w_surface
// KSetWSurface (rk3.cpp:173-190), fill = false: the kinematic surface w from the new u, v (w's halo is
// refreshed next, as rk3.cpp:199)
MI_POINT_KERNEL k_set_w_surface(Range r, Args a) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const idx_t x = g.at(i, j, 0), x2 = g.at2(i, j), sj = g.sj, sk = g.sk;
  const double* ht = f.ht;
  const real *u = f.u2, *v = f.v2;
  (void)k;  // one level
  const double ws =
      f.msfty[x2] * 0.5 * c.rdy *
          ((ht[x2 + sj] - ht[x2]) * (c.cf1 * v[x + sj] + c.cf2 * v[x + sj + sk] + c.cf3 * v[x + sj + 2 * sk]) +
           (ht[x2] - ht[x2 - sj]) * (c.cf1 * v[x] + c.cf2 * v[x + sk] + c.cf3 * v[x + 2 * sk])) +
      f.msftx[x2] * 0.5 * c.rdx *
          ((ht[x2 + 1] - ht[x2]) * (c.cf1 * u[x + 1] + c.cf2 * u[x + 1 + sk] + c.cf3 * u[x + 1 + 2 * sk]) +
           (ht[x2] - ht[x2 - 1]) * (c.cf1 * u[x] + c.cf2 * u[x + sk] + c.cf3 * u[x + 2 * sk]));
  put<kHMass>(f.w2, g, i, j, 0, ws);
}

// ----------------------------------------------------------------------------- calc_cq
// KCq (reference src/dyn/large_step.cpp:127-148): per species, the two neighbours, WRF's order
MI_POINT_KERNEL k_cq(Range r, Args a, Scalars s) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const idx_t x = g.at(i, j, k);
  if (j < g.ny) {
    double t = 0;
    for (int q = 0; q < s.n_moist; ++q) t = t + s.q2[q][x] + s.q2[q][x - 1];
    f.cqu[x] = 1.0 / (1.0 + 0.5 * t);
  }
  if (i < g.nx && j < g.nyv) {
    double t = 0;
    for (int q = 0; q < s.n_moist; ++q) t = t + s.q2[q][x] + s.q2[q][x - g.sj];
    f.cqv[x] = 1.0 / (1.0 + 0.5 * t);
  }
  if (i < g.nx && j < g.ny && k >= 1) {
    double t = 0;
    for (int q = 0; q < s.n_moist; ++q) t = t + s.q2[q][x] + s.q2[q][x - g.sk];
    f.cqw[x] = 0.5 * t;
  }
}

// ----------------------------------------------------------------------------- scalar transport
// The species a launch advances (ScalarTransport::update's batch, scalars.cpp:311-434)
struct SpList {
  int n;
  int sp[MI_MAX_SCALARS];
};

// Face fluxes of one species at the faces of cell (i, j, k) (reference Fluxes, src/dyn/scalars.cpp:42-140;
// periodic or symmetric edges): q = q_2 (stage value), qo = q_1, with the acoustic-step mass fluxes
// ru_m, rv_m, ww_m and the stage-end column mass muts.  Each face is recomputed by the cells on both
// sides with the same expression, which gives the reference's stored face values bit for bit.  The
// arithmetic is `acc`: double in FP64 (the reference's expressions), float in FP32-bulk.
struct ScFaces {
  const Args& a;
  const real* q;
  const real* qo;
  acc dt;
  bool pd;
  __device__ acc f5(acc m3, acc m2, acc m1, acc q0, acc p1, acc p2, acc v) const {
    return pd ? flux5_pd<acc>(m3, m2, m1, q0, p1, p2, v, acc(a.q.hup)) : flux5<acc>(m3, m2, m1, q0, p1, p2, v, acc(a.q.hup));
  }
  __device__ acc f3(acc m2, acc m1, acc q0, acc p1, acc v, acc up) const {
    return pd ? flux3_pd<acc>(m2, m1, q0, p1, v, up) : flux3<acc>(m2, m1, q0, p1, v, up);
  }
  __device__ bool sp() const { return a.g.spec(); }
  // x face at U point y (west face of cell fi), the order degraded next to a specified edge
  __device__ acc fx(idx_t y, int fi) const {
    const acc vel = a.f.ru_m[y];
    const int o = ord_face(fi, a.g.nx, a.g.flx());
    if (o == 2) return acc(H) * vel * (acc(q[y]) + acc(q[y - 1]));
    if (o == 3) return vel * f3(q[y - 2], q[y - 1], q[y], q[y + 1], vel, acc(a.q.hup));
    return vel * f5(q[y - 3], q[y - 2], q[y - 1], q[y], q[y + 1], q[y + 2], vel);
  }
  __device__ acc fy(idx_t y, int fj) const {
    const idx_t sj = a.g.sj;
    const acc vel = a.f.rv_m[y];
    const int o = ord_face(fj, a.g.ny, a.g.fly());
    if (o == 2) return acc(H) * vel * (acc(q[y]) + acc(q[y - sj]));
    if (o == 3) return vel * f3(q[y - 2 * sj], q[y - sj], q[y], q[y + sj], vel, acc(a.q.hup));
    return vel * f5(q[y - 3 * sj], q[y - 2 * sj], q[y - sj], q[y], q[y + sj], q[y + 2 * sj], vel);
  }
  // the spec zone (scalars.cpp:100-103): not advected
  __device__ bool spec_cell(int i, int j) const {
    return sp() && (i == 0 || i == a.g.nx - 1 || j == 0 || j == a.g.ny - 1);
  }
  // the outermost row of a specified or open boundary: the limiter leaves it alone (scalars.cpp:104-107)
  __device__ bool edge_cell(int i, int j) const {
    return (a.g.flx() && (i == 0 || i == a.g.nx - 1)) || (a.g.fly() && (j == 0 || j == a.g.ny - 1));
  }
  __device__ bool open_col(int i) const { return a.g.ox() && (i == 0 || i == a.g.nx - 1); }
  __device__ bool open_row(int j) const { return a.g.oy() && (j == 0 || j == a.g.ny - 1); }
  // the open terms, x then y (scalars.cpp:110-116); fo: q itself (advect_scalar) or q_1 (advect_scalar_pd)
  __device__ double open_terms(double r, int i, int j, int k, const real* fo) const {
    if (open_col(i)) r = r - open_x_scalar(a.g, q, fo, a.f.ru_m, a.c.rdx, i, j, k);
    if (open_row(j)) r = r - open_y_scalar(a.g, q, fo, a.f.rv_m, a.c.rdy, i, j, k);
    return r;
  }
  // w level kk of the column whose level-0 index is y0
  __device__ acc fz(idx_t y0, int kk) const {
    const int nz = a.g.nz;
    if (kk == 0 || kk == nz) return 0;
    const idx_t sk = a.g.sk, y = y0 + kk * sk;
    const acc vel = a.f.ww_m[y];
    if (kk == 1 || kk == nz - 1) return vel * (acc(a.c.fnm[kk]) * acc(q[y]) + acc(a.c.fnp[kk]) * acc(q[y - sk]));
    return vel * f3(q[y - 2 * sk], q[y - sk], q[y], q[y + sk], -vel, acc(a.q.vup));
  }
  __device__ acc mh(idx_t y2, int k) const { return acc(a.c.c1h[k]) * acc(a.f.muts[y2]) + acc(a.c.c2h[k]); }
  // the upwind fluxes of the limiter: the face width dx = 2 / (m + m) / rdx and dx / dt come precomputed
  // (k_pd_geom: the same expressions, once per face column instead of once per level and species)
  __device__ acc fxl(idx_t y, idx_t y2, int k, int fi) const {
    const Fields& f = a.f;
    const acc dx = acc(f.pd_dx[y2]), mu = acc(H) * (mh(y2, k) + mh(y2 - 1, k)), vel = f.ru_m[y];
    // advect_scalar_pd, degrade_xs, i == ids+1 only: vel = ru/mu, cr = vel*dt/dx (scalars.cpp:88-89)
    const acc cr = (a.g.flx() && fi == 1) ? qdiv(qdiv(vel, mu) * dt, dx) : qdiv(qdiv(vel * dt, dx), mu);
    return mu * acc(f.pd_dxdt[y2]) * flux_upwind<acc>(qo[y - 1], qo[y], cr);
  }
  __device__ acc fyl(idx_t y, idx_t y2, int k) const {
    const Fields& f = a.f;
    const idx_t sj = a.g.sj;
    const acc dy = acc(f.pd_dy[y2]), mu = acc(H) * (mh(y2, k) + mh(y2 - sj, k)), vel = f.rv_m[y];
    const acc cr = qdiv(qdiv(vel * dt, dy), mu);
    return mu * acc(f.pd_dydt[y2]) * flux_upwind<acc>(qo[y - sj], qo[y], cr);
  }
  __device__ acc fzl(idx_t y0, idx_t y2, int kk) const {
    if (kk == 0 || kk == a.g.nz) return 0;
    const idx_t sk = a.g.sk, y = y0 + kk * sk;
    const acc dz = acc(a.f.pd_dz[kk]), mu = mh(y2, kk), vel = a.f.ww_m[y];
    const acc cr = qdiv(qdiv(vel * dt, dz), mu);
    return mu * acc(a.f.pd_dz[a.g.nz + 1 + kk]) * flux_upwind<acc>(qo[y - sk], qo[y], cr);
  }
};

// The face widths of the limiter's upwind fluxes (Fluxes::fxl/fyl/fzl, scalars.cpp:118-138) and their ratio
// to the stage's dt: dx at the x face of (i, j), dy at the y face, dz at w level k (then dz / dt at nz+1+k)
__global__ void k_pd_geom(Args a, double dt) {
  const Geo& g = a.g;
  const Fields& f = a.f;
  const int e = int(blockIdx.x) * 256 + int(threadIdx.x);
  const int w = g.nx + 1, h = g.two_d ? 1 : g.ny + 1;
  if (e < w * h) {
    const int i = e % w, j = e / w;
    const idx_t y2 = g.at2(i, j);
    const double dx = 2.0 / (f.msfty[y2] + f.msfty[y2 - 1]) / a.c.rdx;
    f.pd_dx[y2] = dx;
    f.pd_dxdt[y2] = dx / dt;
    const double dy = 2.0 / (f.msftx[y2] + f.msftx[y2 - g.sj]) / a.c.rdy;
    f.pd_dy[y2] = dy;
    f.pd_dydt[y2] = dy / dt;
    return;
  }
  const int kk = e - w * h;
  if (kk >= 1 && kk < g.nz) {
    const double dz = 2.0 / (a.c.rdnw[kk] + a.c.rdnw[kk - 1]);
    f.pd_dz[kk] = dz;
    f.pd_dz[g.nz + 1 + kk] = dz / dt;
  }
}

__device__ __forceinline__ acc m_full(const Lv& c, int k, double mub, double mu) {
  return acc(c.c1h[k]) * (acc(mub) + acc(mu)) + acc(c.c2h[k]);
}

// rk_update_scalar (KUpdate, scalars.cpp:145-158): advect_scalar's tendency and the mass-weighted update,
// every species of the list; the new values go to the species' second buffer
template <bool OPEN, bool QVF>
MI_POINT_KERNEL k_sc_update(Range r, Args a, Scalars s, SpList l, double dt_rk) {
  MI_POINT_INDEX_SP(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j), x0 = g.at(i, j, 0);
  const acc msftx = acc(f.msftx[x2]), msfty = acc(f.msfty[x2]);
  const acc m_old = m_full(c, k, f.mub[x2], f.mu1[x2]), m_new = m_full(c, k, f.mub[x2], f.mu2[x2]);
  {
    const int n = l.sp[sq_];
    const ScFaces fl{a, s.q2[n], s.q1[n], acc(dt_rk), false};
    acc adv = 0;
    if ((OPEN && fl.open_col(i)) || (OPEN && fl.open_row(j))) {  // WRF order: y flux, x flux, open x, open y, vertical (scalars.cpp:129-134)
      acc r0 = 0;
      if (!g.two_d && !(OPEN && fl.open_row(j))) r0 = r0 - (msftx * acc(c.rdy)) * (fl.fy(x + g.sj, j + 1) - fl.fy(x, j));
      if (!(OPEN && fl.open_col(i))) r0 = r0 - (msftx * acc(c.rdx)) * (fl.fx(x + 1, i + 1) - fl.fx(x, i));
      r0 = acc(fl.open_terms(double(r0), i, j, k, s.q2[n]));
      adv = r0 - acc(c.rdnw[k]) * (fl.fz(x0, k + 1) - fl.fz(x0, k));
    } else if (!fl.spec_cell(i, j)) {
      adv = -(msftx * acc(c.rdx)) * (fl.fx(x + 1, i + 1) - fl.fx(x, i));
      if (!g.two_d) adv -= (msftx * acc(c.rdy)) * (fl.fy(x + g.sj, j + 1) - fl.fy(x, j));
      adv = adv - acc(c.rdnw[k]) * (fl.fz(x0, k + 1) - fl.fz(x0, k));
    }
    acc tend = adv * msfty;
    if (QVF && f.qvf && n == 0) {  // set_tend(RQVFTEN) of the vapour (scalars.cpp:147-150)
      if (fl.spec_cell(i, j)) {  // adv_spec (scalars.cpp:119-125): no x flux in the outer columns, no y in the rows
        const bool scol = i == 0 || i == g.nx - 1, srow = j == 0 || j == g.ny - 1;
        acc rs = 0;
        if (!scol) rs = -(msftx * acc(c.rdx)) * (fl.fx(x + 1, i + 1) - fl.fx(x, i));
        if (!g.two_d && !srow) rs -= (msftx * acc(c.rdy)) * (fl.fy(x + g.sj, j + 1) - fl.fy(x, j));
        rs = rs - acc(c.rdnw[k]) * (fl.fz(x0, k + 1) - fl.fz(x0, k));
        f.qvf[x] = rs * msfty;
      } else {
        f.qvf[x] = tend;
      }
    }
    tend += acc((a.q.use_tendf ? s.qtw[n][x] : real(0)));
    put<kHMass>(s.qn[n], g, i, j, k * g.sk, qdiv(m_old * acc(s.q1[n][x]) + acc(dt_rk) * tend, m_new));
  }
}

// rk_update_scalar_pd, first step (KPreUpdatePD, scalars.cpp:174-181): the held tendency folded into
// q_1 (q_1's halo is refreshed next, as scalars.cpp:383)
MI_POINT_KERNEL k_sc_pd_pre(Range r, Args a, Scalars s, SpList l, double dt_rk) {
  MI_POINT_INDEX_SP(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const acc m = m_full(a.c, k, f.mub[x2], f.mu1[x2]);
  {
    const int n = l.sp[sq_];
    put<kHMass>(s.q1[n], g, i, j, k * g.sk,
                qdiv(m * acc(s.q1[n][x]) + acc(dt_rk) * acc((a.q.use_tendf ? s.qtw[n][x] : real(0))), m));
  }
}

// advect_scalar_pd's high-minus-low and low-order fluxes of the six faces of a cell (KFacePDX/Y/Z,
// scalars.cpp:186-212)
struct PdFaces {
  acc dhx0, lox0, dhx1, lox1, dhy0, loy0, dhy1, loy1, dhz0, loz0, dhz1, loz1;
};
__device__ __forceinline__ PdFaces pd_faces(const ScFaces& fl, const Geo& g, int i, int j, int k) {
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j), x0 = g.at(i, j, 0);
  const acc lox0 = fl.fxl(x, x2, k, i);
  const acc dhx0 = fl.fx(x, i) - lox0;
  const acc lox1 = fl.fxl(x + 1, x2 + 1, k, i + 1);
  const acc dhx1 = fl.fx(x + 1, i + 1) - lox1;
  acc loy0 = 0, dhy0 = 0, loy1 = 0, dhy1 = 0;
  if (!g.two_d) {
    loy0 = fl.fyl(x, x2, k);
    dhy0 = fl.fy(x, j) - loy0;
    loy1 = fl.fyl(x + g.sj, x2 + g.sj, k);
    dhy1 = fl.fy(x + g.sj, j + 1) - loy1;
  }
  const acc loz0 = fl.fzl(x0, x2, k);
  const acc dhz0 = fl.fz(x0, k) - loz0;
  const acc loz1 = fl.fzl(x0, x2, k + 1);
  const acc dhz1 = fl.fz(x0, k + 1) - loz1;
  return PdFaces{dhx0, lox0, dhx1, lox1, dhy0, loy0, dhy1, loy1, dhz0, loz0, dhz1, loz1};
}

// advect_scalar_pd's limiter (KScalePD, scalars.cpp:220-248): the scale of each cell's outgoing
// high-minus-low fluxes (its halo is refreshed next, as scalars.cpp:411)
MI_POINT_KERNEL k_sc_pd_scale(Range r, Args a, Scalars s, SpList l, double dt_rk) {
  MI_POINT_INDEX_SP(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const acc msftx = acc(f.msftx[x2]), msfty = acc(f.msfty[x2]);
  const acc mxy = msftx * msfty;
  const acc dt = acc(dt_rk), zero = 0;
  {
    const int n = l.sp[sq_];
    const ScFaces fl{a, s.q2[n], s.q1[n], dt, true};
    if (fl.edge_cell(i, j)) {  // WRF limits interior cells only
      put<kHMass>(s.scale[n], g, i, j, k * g.sk, 1.0);
      return;
    }
    const PdFaces p = pd_faces(fl, g, i, j, k);
    acc hlow = acc(c.rdx) * (p.lox1 - p.lox0);
    if (!g.two_d) hlow += acc(c.rdy) * (p.loy1 - p.loy0);
    const acc ph_low = ((acc(c.c1h[k]) * acc(f.mub[x2]) + acc(c.c2h[k])) + acc(c.c1h[k]) * acc(f.mu1[x2])) * acc(s.q1[n][x]) -
                       dt * (mxy * hlow + msfty * acc(c.rdnw[k]) * (p.loz1 - p.loz0));
    acc hout = acc(c.rdx) * ((p.dhx1 > zero ? p.dhx1 : zero) - (p.dhx0 < zero ? p.dhx0 : zero));
    if (!g.two_d) hout += acc(c.rdy) * ((p.dhy1 > zero ? p.dhy1 : zero) - (p.dhy0 < zero ? p.dhy0 : zero));
    const acc out = mxy * hout + msfty * acc(c.rdnw[k]) * ((p.dhz1 < zero ? p.dhz1 : zero) - (p.dhz0 > zero ? p.dhz0 : zero));
    const acc flux_out = dt * out;
    acc sc = 1;
    if (flux_out > ph_low) {
      sc = qdiv(ph_low, flux_out + acc(1e-20));
      if (sc < zero) sc = 0;
    }
    put<kHMass>(s.scale[n], g, i, j, k * g.sk, sc);
  }
}

// advect_scalar_pd's limited divergence and rk_update_scalar (KUpdatePD, scalars.cpp:252-291)
template <bool OPEN, bool QVF>
MI_POINT_KERNEL k_sc_pd_update(Range r, Args a, Scalars s, SpList l, double dt_rk) {
  MI_POINT_INDEX_SP(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int nz = g.nz;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j), sj = g.sj, sk = g.sk;
  const acc msftx = acc(f.msftx[x2]), msfty = acc(f.msfty[x2]);
  const acc m_old = m_full(c, k, f.mub[x2], f.mu1[x2]), m_new = m_full(c, k, f.mub[x2], f.mu2[x2]);
  const acc dt = acc(dt_rk);
  {
    const int n = l.sp[sq_];
    const real* sc = s.scale[n];
    const ScFaces fl{a, s.q2[n], s.q1[n], dt, true};
    const bool ften = QVF && f.qvf && n == 0;
    if (fl.spec_cell(i, j) && !ften) {  // the spec zone: only the held tendency, folded into q_1 before
      put<kHMass>(s.qn[n], g, i, j, k * sk, qdiv(m_old * acc(s.q1[n][x]), m_new));
      return;
    }
    const PdFaces p = pd_faces(fl, g, i, j, k);
    // the limited flux of a face: scaled by the scale of the cell it leaves
    auto lim = [](acc d, acc s_minus, acc s_plus) { return d * (d > acc(0) ? s_minus : (d < acc(0) ? s_plus : acc(1))); };
    const acc sc0 = sc[x];
    const acc dx0 = lim(p.dhx0, sc[x - 1], sc0), dx1 = lim(p.dhx1, sc0, sc[x + 1]);
    // eta decreases upward: a positive w-level flux leaves the cell above it
    const acc dz0 = k == 0 ? acc(0) : lim(p.dhz0, sc0, sc[x - sk]);
    const acc dz1 = k + 1 == nz ? acc(0) : lim(p.dhz1, sc[x + sk], sc0);
    if (fl.spec_cell(i, j)) {  // RQVFTEN in the spec zone: WRF still forms the tendency, without the edge's own
      // flux (scalars.cpp:262-269); the value is only the held tendency's
      const bool scol = i == 0 || i == g.nx - 1, srow = j == 0 || j == g.ny - 1;
      acc t = -acc(c.rdnw[k]) * (dz1 - dz0 + p.loz1 - p.loz0);
      if (!scol) t -= msftx * (acc(c.rdx) * (dx1 - dx0 + p.lox1 - p.lox0));
      if (!g.two_d && !srow) {
        const acc dy0 = lim(p.dhy0, sc[x - sj], sc0), dy1 = lim(p.dhy1, sc0, sc[x + sj]);
        t -= msftx * (acc(c.rdy) * (dy1 - dy0 + p.loy1 - p.loy0));
      }
      f.qvf[x] = t * msfty;
      put<kHMass>(s.qn[n], g, i, j, k * sk, qdiv(m_old * acc(s.q1[n][x]), m_new));
      return;
    }
    acc tend;
    if ((OPEN && fl.open_col(i)) || (OPEN && fl.open_row(j))) {  // open: open terms (with q_1), vertical, x (not outer columns), y
      tend = acc(fl.open_terms(0.0, i, j, k, s.q1[n]));
      tend = tend - acc(c.rdnw[k]) * (dz1 - dz0 + p.loz1 - p.loz0);
      if (!(OPEN && fl.open_col(i))) tend -= msftx * (acc(c.rdx) * (dx1 - dx0 + p.lox1 - p.lox0));
      if (!g.two_d && !(OPEN && fl.open_row(j))) {
        const acc dy0 = lim(p.dhy0, sc[x - sj], sc0), dy1 = lim(p.dhy1, sc0, sc[x + sj]);
        tend -= msftx * (acc(c.rdy) * (dy1 - dy0 + p.loy1 - p.loy0));
      }
    } else {
      tend = -acc(c.rdnw[k]) * (dz1 - dz0 + p.loz1 - p.loz0);
      tend -= msftx * (acc(c.rdx) * (dx1 - dx0 + p.lox1 - p.lox0));
      if (!g.two_d) {
        const acc dy0 = lim(p.dhy0, sc[x - sj], sc0), dy1 = lim(p.dhy1, sc0, sc[x + sj]);
        tend -= msftx * (acc(c.rdy) * (dy1 - dy0 + p.loy1 - p.loy0));
      }
    }
    if (ften) f.qvf[x] = tend * msfty;
    put<kHMass>(s.qn[n], g, i, j, k * sk, qdiv(m_old * acc(s.q1[n][x]) + dt * (tend * msfty), m_new));
  }
}

// ----------------------------------------------------------------------------- specified lateral boundaries
// The ring of LateralBC (reference src/dyn/lateral_bc.cpp:34-58): side 0..3 = S, N, W, E; the S and N rows
// b cover i = b .. ibe-b, the W and E columns j = b+1 .. jbe-b-1 (the corners belong to the rows); ibe, jbe
// the last index of the staggering.  Slabs are stored xs, xe, ys, ye: S, N, W, E -> ys, ye, xs, xe.
__device__ __forceinline__ bool ring_map(int pos, int side, int b, int ibe, int jbe, int& i, int& j) {
  switch (side) {
    case 0: j = b; i = pos; return pos >= b && pos <= ibe - b;
    case 1: j = jbe - b; i = pos; return pos >= b && pos <= ibe - b;
    case 2: i = b; j = pos; return pos >= b + 1 && pos <= jbe - b - 1;
    default: i = ibe - b; j = pos; return pos >= b + 1 && pos <= jbe - b - 1;
  }
}
__device__ __forceinline__ int slab_of_side(int side) { return side < 2 ? side + 2 : side - 2; }
__device__ __forceinline__ int bdy_nk(int kind, const Geo& g) { return kind == kBdyPH ? g.nz + 1 : (kind == kBdyMU ? 1 : g.nz); }
__device__ __forceinline__ int bdy_npos(int kind, int sl, const Geo& g) {
  return sl < 2 ? g.ny + (kind == kBdyV ? 1 : 0) : g.nx + (kind == kBdyU ? 1 : 0);
}
// offset of slab sl (0..3 values, 4..7 tendencies) in a variable's buffer
__device__ __forceinline__ int64_t bdy_slab(int kind, int sl, const Geo& g) {
  const int64_t nk = bdy_nk(kind, g), w = 5;
  const int64_t sx = int64_t(bdy_npos(kind, 0, g)) * nk * w, sy = int64_t(bdy_npos(kind, 2, g)) * nk * w;
  const int h = sl / 4, s4 = sl % 4;
  return int64_t(h) * (2 * sx + 2 * sy) + (s4 < 2 ? s4 * sx : 2 * sx + (s4 - 2) * sy);
}

// One launch over the rows of every variable in bs: thread (pos, k), block z = (variable, side, row).
//   kBdySpec   spec_bdytend (KSpecTend, lateral_bc.cpp:67-79): the spec zone's tendency = the data's
//   kBdyRelax  relax_bdytend (KRelax, lateral_bc.cpp:82-121) of rows 1..3 into the held tendencies
//   kBdyFinal  spec_bdy_final (KFinal, lateral_bc.cpp:180-194): the spec zone's values at t + dt (the
//              reference refreshes no halo after it)
__global__ void k_bdy(Args a, Scalars s, BdySet bs, int mode, int b0, int nb) {
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const int pos = int(blockIdx.x) * 64 + int(threadIdx.x), k = int(blockIdx.y), z = int(blockIdx.z);
  const int vi = z / (4 * nb), side = (z / nb) % 4, b = b0 + z % nb;
  if (vi >= bs.n) return;
  const BdyVar V = bs.v[vi];
  const int kind = V.kind, nk = bdy_nk(kind, g);
  if (k >= nk) return;
  const int ibe = kind == kBdyU ? g.nx : g.nx - 1, jbe = kind == kBdyV ? g.ny : g.ny - 1;
  int i, j;
  if (!ring_map(pos, side, b, ibe, jbe, i, j)) return;
  const int sl = slab_of_side(side), npos = bdy_npos(kind, sl, g);
  const double* vals = V.base + bdy_slab(kind, sl, g);
  const double* tens = V.base + bdy_slab(kind, sl + 4, g);
  auto at = [&](int p, int bb) { return int64_t(p) + int64_t(k) * npos + int64_t(bb) * npos * nk; };
  const double dtbc = *f.bdy_time;
  auto bval = [&](int p, int bb) { return vals[at(p, bb)] + dtbc * tens[at(p, bb)]; };
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  if (mode == kBdySpec) {
    const double t = tens[at(pos, b)];
    switch (kind) {
      case kBdyU: f.ru_tend[x] = real(t); break;
      case kBdyV: f.rv_tend[x] = real(t); break;
      case kBdyT: f.t_tend[x] = real(t); break;
      case kBdyPH: f.ph_tend[x] = real(t); break;
      case kBdyMU: f.mu_tend[x2] = t; break;
      default: s.qtw[V.q][x] = real(t); break;
    }
    return;
  }
  if (mode == kBdyFinal) {
    const double bv = bval(pos, b);
    switch (kind) {
      case kBdyU: f.u2[x] = real(f.msfuy[x2] * bv / (c.c1h[k] * f.muus[x2] + c.c2h[k])); break;
      case kBdyV: f.v2[x] = real(f.msfvx[x2] * bv / (c.c1h[k] * f.muvs[x2] + c.c2h[k])); break;
      case kBdyT: f.t2[x] = real(bv / (c.c1h[k] * f.muts[x2] + c.c2h[k])); break;
      case kBdyPH: f.ph2[x] = real(bv / (c.c1f[k] * f.muts[x2] + c.c2f[k])); break;
      case kBdyMU: f.mu2[x2] = bv; break;
      default: s.q2[V.q][x] = real(bv / (c.c1h[k] * f.muts[x2] + c.c2h[k])); break;
    }
    return;
  }
  // the model field as the data holds it, at (ii, jj)
  auto F = [&](int ii, int jj) -> double {
    const idx_t y = g.at(ii, jj, k), y2 = g.at2(ii, jj);
    switch (kind) {
      case kBdyU: return f.ru[y];
      case kBdyV: return f.rv[y];
      case kBdyT: return f.t2[y] * (c.c1h[k] * f.mut[y2] + c.c2h[k]);
      case kBdyPH: return f.ph2[y] * (c.c1f[k] * f.mut[y2] + c.c2f[k]);
      case kBdyMU: return f.mu2[y2];
      default: return s.q2[V.q][y] * (c.c1h[k] * f.mut[y2] + c.c2h[k]);
    }
  };
  double fls0, fls1, fls2, fls3, fls4;
  fls0 = bval(pos, b) - F(i, j);
  if (side <= 1) {
    const int im1 = i - 1 > 0 ? i - 1 : 0, ip1 = i + 1 < ibe ? i + 1 : ibe;
    const int jin = side == 0 ? j + 1 : j - 1, jout = side == 0 ? j - 1 : j + 1;
    fls1 = bval(im1, b) - F(im1, j);
    fls2 = bval(ip1, b) - F(ip1, j);
    fls3 = bval(i, b - 1) - F(i, jout);
    fls4 = bval(i, b + 1) - F(i, jin);
  } else {
    const int iin = side == 2 ? i + 1 : i - 1, iout = side == 2 ? i - 1 : i + 1;
    fls1 = bval(j - 1, b) - F(i, j - 1);
    fls2 = bval(j + 1, b) - F(i, j + 1);
    fls3 = bval(j, b - 1) - F(iout, j);
    fls4 = bval(j, b + 1) - F(iin, j);
  }
  const double fc = a.q.fcx[b], gc = a.q.gcx[b];
  const double sm = fls1 + fls2 + fls3 + fls4 - 4.0 * fls0;
  const double val = fc * fls0 - gc * sm;
  switch (kind) {  // *_save -> *_tendf (u, v times msfuy, msfvx); scalars straight into the tendency
    case kBdyU: f.ru_tendf[x] = real(f.ru_tendf[x] + val * f.msfuy[x2]); break;
    case kBdyV: f.rv_tendf[x] = real(f.rv_tendf[x] + val * f.msfvx[x2]); break;
    case kBdyT: f.t_tendf[x] = real(f.t_tendf[x] + val); break;
    case kBdyPH: f.ph_tendf[x] = real(f.ph_tendf[x] + val); break;
    case kBdyMU: f.mu_tend[x2] = f.mu_tend[x2] + val; break;
    default: s.qtw[V.q][x] = real(s.qtw[V.q][x] + fc * fls0 - gc * sm); break;
  }
}

// zero_grad_bdy on w (KZeroGrad, lateral_bc.cpp:146-161): the spec zone's w from the first row in, every
// level.  The reference runs it after every substep's advance_w; nothing reads the spec zone's w in between
// (advance_w skips those columns), so the copy after the last substep leaves the same values.
__global__ void k_bdy_zero_grad(Args a) {
  const Geo& g = a.g;
  const int pos = int(blockIdx.x) * 64 + int(threadIdx.x), k = int(blockIdx.y), side = int(blockIdx.z);
  const int ibe = g.nx - 1, jbe = g.ny - 1;
  int i, j;
  if (k > g.nz || !ring_map(pos, side, 0, ibe, jbe, i, j)) return;
  const int ic = i < 1 ? 1 : (i > ibe - 1 ? ibe - 1 : i), jc = j < 1 ? 1 : (j > jbe - 1 ? jbe - 1 : j);
  real* w = a.f.w2;
  switch (side) {
    case 0: w[g.at(i, j, k)] = w[g.at(ic, 1, k)]; break;
    case 1: w[g.at(i, j, k)] = w[g.at(ic, jbe - 1, k)]; break;
    case 2: w[g.at(i, j, k)] = w[g.at(1, jc, k)]; break;
    default: w[g.at(i, j, k)] = w[g.at(ibe - 1, jc, k)]; break;
  }
}

// flow_dep_bdy (KFlowDep, lateral_bc.cpp:163-178) of the species without boundary data: the spec zone takes
// the next row in at outflow and 0 at inflow (by the sign of ru_m, rv_m); q_2's halo is refreshed next
__global__ void k_bdy_flow_dep(Args a, Scalars s, SpList l) {
  const Geo& g = a.g;
  const Fields& f = a.f;
  const int pos = int(blockIdx.x) * 64 + int(threadIdx.x), k = int(blockIdx.y);
  const int q = int(blockIdx.z) / 4, side = int(blockIdx.z) % 4;
  const int ibe = g.nx - 1, jbe = g.ny - 1;
  int i, j;
  if (q >= l.n || k >= g.nz || !ring_map(pos, side, 0, ibe, jbe, i, j)) return;
  real* fq = s.q2[l.sp[q]];
  const int ic = i < 1 ? 1 : (i > ibe - 1 ? ibe - 1 : i), jc = j < 1 ? 1 : (j > jbe - 1 ? jbe - 1 : j);
  double v;
  switch (side) {
    case 0: v = f.rv_m[g.at(i, j, k)] < 0.0 ? double(fq[g.at(ic, 1, k)]) : 0.0; break;
    case 1: v = f.rv_m[g.at(i, j + 1, k)] > 0.0 ? double(fq[g.at(ic, jbe - 1, k)]) : 0.0; break;
    case 2: v = f.ru_m[g.at(i, j, k)] < 0.0 ? double(fq[g.at(1, jc, k)]) : 0.0; break;
    default: v = f.ru_m[g.at(i + 1, j, k)] > 0.0 ? double(fq[g.at(ibe - 1, jc, k)]) : 0.0; break;
  }
  put<kHMass>(fq, g, i, j, k * g.sk, v);
}

// LateralBC::advance_time (rk3.cpp:93): dtbc += dt at the start of each step, on the device
__global__ void k_bdy_time(double* t, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *t = *t + dt;
}

// RK3::step's q_1 = q_2 copies (rk3.cpp:99), whole allocations, one species per grid row
__global__ void k_copy_scalars(Scalars s, int64_t n3, int64_t off) {
  const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (e >= n3) return;
  const int n = int(blockIdx.y);
  (s.q1[n] - off)[e] = (s.q2[n] - off)[e];
}

// LargeStep::cfl_max's pointwise Courant numbers (FVertCfl, FHorizCfl; large_step.cpp:672-687) over levels
// 1..nz-1, reduced with an atomic max of the bit patterns: for non-negative doubles the integer order is the
// numeric order, so the maximum is exact and independent of the order of the points; NaN counts as 0
MI_POINT_KERNEL k_cfl(Range r, Args a, unsigned long long* out) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const Lv& c = a.c;
  const idx_t x = g.at(i, j, k), x2 = g.at2(i, j);
  const double dt = a.q.dt;
  double v = f.ww[x] / (c.c1f[k] * f.mut[x2] + c.c2f[k]) * c.rdnw[k] * dt;
  v = v < 0.0 ? -v : v;
  double ha = f.u2[x] * c.rdx * f.msfux[x2] * dt, hb = f.v2[x] * c.rdy * f.msfvy[x2] * dt;
  ha = ha < 0.0 ? -ha : ha;
  hb = hb < 0.0 ? -hb : hb;
  const double h = hb > ha ? hb : ha;
  if (v > 0.0) atomicMax(&out[0], static_cast<unsigned long long>(__double_as_longlong(v)));
  if (h > 0.0) atomicMax(&out[1], static_cast<unsigned long long>(__double_as_longlong(h)));
}

// finiteness of the prognostic fields (flags[fld] = 1 when a NaN or an infinity is found)
MI_POINT_KERNEL k_finite(Range r, Args a, int* flags) {
  MI_POINT_INDEX(r)
  const Geo& g = a.g;
  const Fields& f = a.f;
  const idx_t x = g.at(i, j, k);
  auto bad = [](double v) { return !(v - v == 0.0); };
  if (i < g.nx + 1 && j < g.ny && k < g.nz && bad(f.u2[x])) flags[MI_U] = 1;
  if (i < g.nx && j < g.nyv && k < g.nz && bad(f.v2[x])) flags[MI_V] = 1;
  if (i < g.nx && j < g.ny) {
    if (bad(f.w2[x])) flags[MI_W] = 1;
    if (bad(f.ph2[x])) flags[MI_PH] = 1;
    if (k < g.nz && bad(f.t2[x])) flags[MI_T] = 1;
    if (k == 0 && bad(f.mu2[g.at2(i, j)])) flags[MI_MU] = 1;
  }
}

}  // namespace

// ============================================================================= launchers
void launch_halo(const Args& a, const HaloSet& s, hipStream_t st) {
  if (s.n == 0) return;
  const Geo& g = a.g;
  int nymax = 0, nxmax = 0;
  bool odd_face = false;
  for (int f = 0; f < s.n; ++f) {
    nymax = s.nys[f] > nymax ? s.nys[f] : nymax;
    nxmax = s.nxs[f] > nxmax ? s.nxs[f] : nxmax;
    odd_face = odd_face || (g.bcx == MI_BC_SYMMETRIC && s.px[f] == 2 && s.face_x[f]) ||
               (g.bcy == MI_BC_SYMMETRIC && s.py[f] == 2 && s.face_y[f]);
  }
  const unsigned nfk = unsigned(s.n * s.nkmax);
  const bool t = timing("halo");
  if (t) time_begin(st);
  if (!odd_face) {  // one pass
    const int pts = (nymax + 1) * 8 + (g.two_d ? 0 : (nxmax + 1 + 2 * kHalo) * 8);
    hipLaunchKernelGGL(k_halo_xy, dim3(unsigned((pts + 255) / 256), nfk, 1), dim3(256, 1, 1), 0, st, g, s);
    if (t) time_end(st, 0.0);
    return;
  }
  hipLaunchKernelGGL(k_halo_x, dim3(unsigned((nymax + 31) / 32), nfk, 1), dim3(256, 1, 1), 0, st, g, s);
  if (!g.two_d) {
    const int ni = nxmax + 2 * kHalo;
    hipLaunchKernelGGL(k_halo_y, dim3(unsigned((ni + 63) / 64), nfk, 1), dim3(64, 8, 1), 0, st, g, s);
  }
  if (t) time_end(st, 0.0);
}

void launch_copy_step(const Args& a, int64_t n3, int64_t n2, hipStream_t st) {
  // offsets of point (0,0,0) inside the allocations: identical for every field of the domain
  const int64_t off3 = int64_t(kLpad) + int64_t(a.g.hy) * a.g.sj;
  // mu's rows only: rows x pitch values (an x-z slice: its one row of sk values)
  const int64_t pitch = a.g.sk, sj = a.g.sj, rows = sj > 0 ? n2 / sj : 1, n2rows = rows * pitch;
  hipLaunchKernelGGL(k_copy_step, dim3(unsigned((n3 + 255) / 256), 6, 1), dim3(256, 1, 1), 0, st, a.f, nullptr, nullptr, n3,
                     n2rows, off3, off3, off3, off3, off3, off3, sj, pitch);
}

void launch_fill(double* p, int64_t n, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(unsigned((n + 255) / 256), 1, 1), dim3(256, 1, 1), 0, st, p, n, v);
}
void launch_fill2d(const Args& a, double* p, int64_t n_alloc, double v, hipStream_t st) {
  const Geo& g = a.g;
  if (g.sj <= 0) {  // an x-z slice: one row
    launch_fill(p, g.sk, v, st);
    return;
  }
  const int64_t rows = n_alloc / g.sj, n = g.sk * rows;
  hipLaunchKernelGGL(k_fill2d, dim3(unsigned((n + 255) / 256), 1, 1), dim3(256, 1, 1), 0, st, p, int64_t(g.sj),
                     int(g.sk), rows, v);
}

constexpr size_t kLdsMax = 65536;
size_t col_lds_bytes(int ci, int arrays, int nz, int extra) {
  return (size_t(arrays) * size_t(nz + 1) + size_t(extra)) * size_t(ci) * sizeof(double);
}

template <template <int, int> class Launch, class... T>
void dispatch_cfg(int cfg, T&&... args) {
  switch (cfg) {
    case 0: Launch<16, 16>::go(args...); break;
    case 1: Launch<16, 32>::go(args...); break;
    case 2: Launch<32, 16>::go(args...); break;
    case 3: Launch<8, 32>::go(args...); break;
    case 4: Launch<8, 64>::go(args...); break;
    case 5: Launch<4, 128>::go(args...); break;
    case 6: Launch<2, 256>::go(args...); break;
    case 7: Launch<64, 4>::go(args...); break;
    case 8: Launch<64, 8>::go(args...); break;
    case 9: Launch<64, 16>::go(args...); break;
    case 10: Launch<32, 8>::go(args...); break;
    case 11: Launch<16, 64>::go(args...); break;
    default: Launch<4, 64>::go(args...); break;
  }
}

// the points a column kernel advances: the substep kernel leaves the spec zone's columns to k_acoustic_ring
double col_points(const char* name, const Args& a) {
  double cols = double(a.g.nx) * double(a.g.ny);
  if (a.g.bcx == MI_BC_SPECIFIED && std::strcmp(name, "acoustic") == 0) cols -= 2.0 * (a.g.nx + a.g.ny) - 4.0;
  return cols * double(a.g.nz);
}

// one launch of a column kernel with block shape (CI, KT), timed when selected
#define MI_COL_LAUNCHER(Name, kern)                                                                              \
  template <int CI, int KT>                                                                                    \
  struct Name {                                                                                                \
    template <class... T>                                                                                      \
    static void go(const char* name, int arrays, int extra, int ncx, int ncy, hipStream_t st, const Args& a,    \
                   T... args) {                                                                                \
      const size_t lds = col_lds_bytes(CI, arrays, a.g.nz, extra);                                            \
      const bool t = timing(name);                                                                           \
      if (t) time_begin(st);                                                                                 \
      hipLaunchKernelGGL((kern<CI, KT>), dim3(unsigned((ncx + CI - 1) / CI), unsigned(ncy), 1), dim3(CI * KT, 1, 1), \
                         lds, st, a, args...);                                                               \
      if (t) time_end(st, point_bytes(name) * col_points(name, a));                                          \
    }                                                                                                          \
  };
MI_COL_LAUNCHER(LargePrepL, k_large_prep)
MI_COL_LAUNCHER(SsColumnL, k_ss_column)
MI_COL_LAUNCHER(AcousticL, k_acoustic)
MI_COL_LAUNCHER(FinishL, k_finish)

// LDS arrays (per level, extra per column) of each column kernel, in ColKernel order
constexpr int kColArrays[kNumColKernels][2] = {{1, 0}, {3, 0}, {10, 5}, {1, 1}};

bool col_cfg_fits(int kernel, int cfg, int nz) {
  if (cfg == kColMarch) return kernel == kColAcoustic;  // the k-march
  if (cfg < 0 || cfg >= kNumColCfgs) return false;
  if (kernel == kColAcoustic && kColCfgs[cfg][1] < nz + 1) return false;  // one level per thread
  return col_lds_bytes(kColCfgs[cfg][0], kColArrays[kernel][0], nz, kColArrays[kernel][1]) <= kLdsMax;
}

int col_cfg_default(int kernel, int nx, int ny, int nz) {
  // the substep kernel: the narrowest one-level-per-thread shape; the others: one level per thread on
  // small domains (latency), 16 columns (full 128-byte rows) on larger ones
  if (kernel == kColAcoustic) {
    for (int c : {4, 5, 6})
      if (col_cfg_fits(kernel, c, nz)) return c;
    return 6;
  }
  const bool small = int64_t(nx) * ny <= 16384;
  const int pref[] = {small ? (nz + 1 <= 64 ? 4 : (nz + 1 <= 128 ? 5 : 6)) : 0, 3, 5, 6};
  for (int c : pref)
    if (col_cfg_fits(kernel, c, nz)) return c;
  return 6;
}

void launch_large_prep(const Args& a, int cfg, hipStream_t st) {
  dispatch_cfg<LargePrepL>(cfg, "large_prep", 1, 0, a.g.nx + 1, a.g.nyv, st, a);
}

void launch_tendencies(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  const int ui0 = g.bcx != MI_BC_PERIODIC ? 1 : 0;
  const int vj0 = (g.bcy != MI_BC_PERIODIC && !g.two_d) ? 1 : 0;
  // open boundaries: the boundary faces are advanced too (large_step.cpp:804-806)
  // the open-boundary code and the RTHFTEN store compiled in only where used
  const bool op = g.ox() || g.oy();
  launch3("tend_u", op ? k_tend_u<true> : k_tend_u<false>, Range{g.ox() ? 0 : ui0, g.ox() ? g.nx + 1 : g.nx, 0, g.ny, 0, g.nz}, st, a);
  if (!g.two_d)
    launch3("tend_v", op ? k_tend_v<true> : k_tend_v<false>, Range{0, g.nx, g.oy() ? 0 : vj0, g.oy() ? g.ny + 1 : g.ny, 0, g.nz}, st, a);
  launch3("tend_w", op ? k_tend_w<true> : k_tend_w<false>, Range{0, g.nx, 0, g.ny, 1, g.nz + 1}, st, a);
  launch3("tend_t", op ? (a.f.thf ? k_tend_t<true, true> : k_tend_t<true, false>)
                       : (a.f.thf ? k_tend_t<false, true> : k_tend_t<false, false>),
          Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a);
  launch3("tend_ph", op ? k_tend_ph<true> : k_tend_ph<false>, Range{0, g.nx, 0, g.ny, 1, g.nz + 1}, st, a);
}

void launch_ss_prep(const Args& a, bool first, hipStream_t st) {
  const Geo& g = a.g;
  launch3("ss_prep", k_ss_prep, Range{0, g.nx + 1, 0, g.nyv, 0, g.nz + 1}, st, a, first ? 1 : 0);
}
void launch_ss_column(const Args& a, int cfg, double cof, double dts, hipStream_t st) {
  dispatch_cfg<SsColumnL>(cfg, "ss_column", 3, 0, a.g.nx, a.g.ny, st, a, cof, dts);
}
void launch_uv(const Args& a, double dts, hipStream_t st) {
  const Geo& g = a.g;
  // the east (north) faces of the spec zone or of an open direction
  const int ex = (g.spec() || g.ox()) ? 1 : 0, ey = (g.spec() || g.oy()) ? 1 : 0;
  launch3("uv", (g.ox() || g.oy()) ? k_uv<true> : k_uv<false>, Range{0, g.nx + ex, 0, g.ny + ey, 0, g.nz}, st, a, dts);
}
// the batched substep shapes: uv blocks of 64 x 4 points, acoustic blocks of 8 columns x 64 levels
constexpr int kBatchCI = 8, kBatchKT = 64;
int batch_table_uv(const Args* const* as, const double* dts, int n, BatchMember* out) {
  int start = 0;
  for (int m = 0; m < n; ++m) {
    const Geo& g = as[m]->g;
    const int ex = (g.spec() || g.ox()) ? 1 : 0, ey = (g.spec() || g.oy()) ? 1 : 0;
    BatchMember& t = out[m];
    t.i0 = 0; t.i1 = g.nx + ex; t.j0 = 0; t.j1 = g.ny + ey; t.k0 = 0; t.k1 = g.nz;
    t.gx = (t.i1 - t.i0 + kBx - 1) / kBx;
    t.gy = (t.j1 - t.j0 + 3) / 4;
    t.start = start;
    t.dts = dts[m];
    start += t.gx * t.gy * (t.k1 - t.k0);
  }
  out[n].start = start;
  return start;
}
int batch_table_acoustic(const Args* const* as, const double* dts, int n, BatchMember* out) {
  int start = 0;
  for (int m = 0; m < n; ++m) {
    const Geo& g = as[m]->g;
    BatchMember& t = out[m];
    t.gx = (g.nx + kBatchCI - 1) / kBatchCI;
    t.gy = g.ny;
    t.i0 = t.j0 = t.k0 = 0;
    t.i1 = g.nx; t.j1 = g.ny; t.k1 = g.nz;
    t.start = start;
    t.dts = dts[m];
    start += t.gx * t.gy;
  }
  out[n].start = start;
  return start;
}
bool batch_fits(const Args& a) { return a.g.nz + 1 <= kBatchKT; }
void launch_uv_batch(const Args* dev_args, const BatchMember* dev_tab, int n, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_uv_batch, dim3(unsigned(blocks), 1, 1), dim3(kBx, 4, 1), 0, st, dev_args, dev_tab, n);
}
void launch_acoustic_batch(const Args* dev_args, const BatchMember* dev_tab, int n, int blocks, int nzmax, int it, int ns,
                           hipStream_t st) {
  const size_t lds = col_lds_bytes(kBatchCI, kColArrays[kColAcoustic][0], nzmax, kColArrays[kColAcoustic][1]);
  hipLaunchKernelGGL((k_acoustic_batch<kBatchCI, kBatchKT>), dim3(unsigned(blocks), 1, 1), dim3(kBatchCI * kBatchKT, 1, 1),
                     lds, st, dev_args, dev_tab, n, it, ns);
}
// KSumEdge of the open directions after a substep's column kernel
void launch_open_sums(const Args& a, int it, int ns, hipStream_t st) {
  const Geo& g = a.g;
  if (g.ox()) launch3("open_sum", k_open_sum, Range{g.nx - 1, g.nx, 0, g.ny, 0, g.nz}, st, a, it, ns, 0);
  if (g.oy()) launch3("open_sum", k_open_sum, Range{0, g.nx, g.ny - 1, g.ny, 0, g.nz}, st, a, it, ns, 1);
}

void launch_acoustic(const Args& a, int cfg, double dts, int it, int ns, hipStream_t st) {
  if (cfg == kColMarch) {  // one thread per column, 64 columns per block
    const bool t = timing("acoustic");
    if (t) time_begin(st);
    hipLaunchKernelGGL(k_acoustic_march, dim3(unsigned((a.g.nx + 63) / 64), unsigned(a.g.ny), 1), dim3(64, 1, 1), 0, st,
                       a, dts, it, ns);
    if (t) time_end(st, point_bytes("acoustic") * col_points("acoustic", a));
  } else {
    dispatch_cfg<AcousticL>(cfg, "acoustic", 10, 5, a.g.nx, a.g.ny, st, a, dts, it, ns);
  }
  launch_open_sums(a, it, ns, st);
}
void launch_finish(const Args& a, int cfg, bool prhophi, double hd_dt, hipStream_t st) {
  dispatch_cfg<FinishL>(cfg, "finish", 1, 1, a.g.nx + 1, a.g.nyv, st, a, prhophi ? 1 : 0, hd_dt);
}
void launch_sixth_dyn(const Args& a, int mode, hipStream_t st) {
  const Geo& g = a.g;
  const double coef = a.q.diff_6th_factor * 0.015625 / (2.0 * a.q.dt);
  launch3a("sixth", 8, k_sixth_dyn, Range{0, g.nx + 1, 0, g.nyv, 0, g.nz + 1}, st, a, coef, mode);  // u v w t -> 4 tendf
}
void launch_smag2d(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  // defor12 over the vorticity points (u v ph -> defor12), then K (u v ph defor12 -> xkmh xkhh)
  launch3a("deform12", 4, k_deform12, Range{0, g.nx + 1, 0, g.two_d ? 1 : g.ny + 1, 0, g.nz}, st, a);
  const bool fx = g.ox() || g.spec(), fy = g.oy() || g.spec();
  int is = fx ? 1 : 0, ie = fx ? g.nx - 1 : g.nx;
  if (g.bcx == MI_BC_PERIODIC) {
    is = 0;
    ie = g.nx;
  }
  launch3a("smag2d", 6, k_smag2d, Range{is, ie, fy ? 1 : 0, fy ? g.ny - 1 : g.ny, 0, g.nz}, st, a);
}
void launch_dampkm(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  int is = 0, ie = g.nx, js = 0, je = g.ny;
  if (g.spec()) {  // one row in from the relaxation zone (spec_bdy_width 5)
    is = std::max(is, MI_SPEC_BDY_WIDTH - 1);
    ie = std::min(ie, g.nx - MI_SPEC_BDY_WIDTH + 1);
    js = std::max(js, MI_SPEC_BDY_WIDTH - 1);
    je = std::min(je, g.ny - MI_SPEC_BDY_WIDTH + 1);
  }
  launch3a("dampkm", 3, k_dampkm, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a, is, ie, js, je);  // ph -> xkmh xkhh
}
void launch_rayleigh(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  // u v w t ph -> their 4 held tendencies (read and written)
  launch3a("rayleigh", 9, k_rayleigh, Range{0, g.nx + 1, 0, g.nyv, 0, g.nz + 1}, st, a);
}
void launch_sixth_scalars(const Args& a, const Scalars& s, int n, double dt_rk, hipStream_t st) {
  if (n <= 0) return;
  const Geo& g = a.g;
  const double coef = a.q.diff_6th_factor * 0.015625 / (2.0 * dt_rk);
  launch3a("sixth_sc", 3 * n, k_sixth_sc, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a, s, n, coef);  // q, tendf -> tendf
}
void launch_set_w_surface(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  launch3a("set_w_surface", 7, k_set_w_surface, Range{0, g.nx, 0, g.ny, 0, 1}, st, a);  // u, v (3 levels each) -> w (1)
}
void launch_cfl(const Args& a, unsigned long long* out, hipStream_t st) {
  const Geo& g = a.g;
  launch3a("cfl", 4, k_cfl, Range{0, g.nx, 0, g.ny, 1, g.nz}, st, a, out);  // ww u v (mut 2-D) read
}

// The scratch of the allocation-free CFL maxima (mi_cfl_max_noalloc): the two running maxima live in this
// file-scope device array, one copy per GPU, held until the process ends.  Host code never uses its name as an
// address: hipGetSymbolAddress gives each device's copy once, with that device current.
static __device__ __align__(256) unsigned long long scratch[kCflScratchSlots];
namespace {
__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int e = int(blockIdx.x) * 64 + int(threadIdx.x);
  if (e < n) p[e] = 0ull;
}
constexpr int kMaxDevices = 64;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

unsigned long long* cfl_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratch_ptr[device] = static_cast<unsigned long long*>(ptr);
  }
  return g_scratch_ptr[device];
}

void launch_cfl_scratch(const Args& a, unsigned long long* sp, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u64, dim3(1, 1, 1), dim3(64, 1, 1), 0, st, sp, kCflScratchSlots);  // a kernel, not a memset
  launch_cfl(a, sp, st);
}
void launch_cq(const Args& a, const Scalars& s, hipStream_t st) {
  const Geo& g = a.g;
  // the moist species -> cqu, cqv, cqw
  launch3a("cq", s.n_moist + 3, k_cq, Range{0, g.nx + 1, 0, g.nyv, 0, g.nz}, st, a, s);
}
// the kernels of the scalar transport of the species sp[0..n) (step: 0 rk_update_scalar; 1, 2, 3 the
// positive-definite pre-update, limiter scale and update); algorithmic arrays: the mass fluxes ru_m, rv_m,
// ww_m once, then per species
void launch_pd_geom(const Args& a, double dt, hipStream_t st) {
  const Geo& g = a.g;
  const int pts = (g.nx + 1) * (g.two_d ? 1 : g.ny + 1) + g.nz + 1;
  hipLaunchKernelGGL(k_pd_geom, dim3(unsigned((pts + 255) / 256), 1, 1), dim3(256, 1, 1), 0, st, a, dt);
}

void launch_scalars(const Args& a, const Scalars& s, const int* sp, int n, int step, double dt_rk, hipStream_t st) {
  if (n <= 0) return;
  SpList l{};
  l.n = n;
  for (int q = 0; q < n; ++q) l.sp[q] = sp[q];
  const Geo& g = a.g;
  const Range r{0, g.nx, 0, g.ny, 0, g.nz};
  // the open-boundary terms and the RQVFTEN store compiled in only where used
  const bool op = g.ox() || g.oy(), qv = a.f.qvf != nullptr;
  switch (step) {
    case 0: launch3s("sc_update", 3 + 4 * n, op ? (qv ? k_sc_update<true, true> : k_sc_update<true, false>)
                                                  : (qv ? k_sc_update<false, true> : k_sc_update<false, false>),
                     r, n, st, a, s, l, dt_rk); break;        // q_2 q_1 q_tendf -> q_new
    case 1: launch3s("sc_pd_pre", 3 * n, k_sc_pd_pre, r, n, st, a, s, l, dt_rk); break;            // q_1 q_tendf -> q_1
    case 2: launch3s("sc_pd_scale", 3 + 3 * n, k_sc_pd_scale, r, n, st, a, s, l, dt_rk); break;    // q_2 q_1 -> scale
    default: launch3s("sc_pd_update", 3 + 4 * n, op ? (qv ? k_sc_pd_update<true, true> : k_sc_pd_update<true, false>)
                                                        : (qv ? k_sc_pd_update<false, true> : k_sc_pd_update<false, false>),
                      r, n, st, a, s, l, dt_rk); break; // q_2 q_1 scale -> q_new
  }
}
void launch_bdy(const Args& a, const Scalars& s, const BdySet& bs, BdyMode mode, hipStream_t st) {
  if (bs.n <= 0) return;
  const Geo& g = a.g;
  const int maxpos = std::max(g.nx, g.ny) + 1;
  const int b0 = mode == kBdyRelax ? 1 : 0, nb = mode == kBdyRelax ? 3 : 1;
  const char* name = mode == kBdySpec ? "bdy_spec" : (mode == kBdyRelax ? "bdy_relax" : "bdy_final");
  const bool t = timing(name);
  if (t) time_begin(st);
  hipLaunchKernelGGL(k_bdy, dim3(unsigned((maxpos + 63) / 64), unsigned(g.nz + 1), unsigned(bs.n * 4 * nb)), dim3(64, 1, 1),
                     0, st, a, s, bs, int(mode), b0, nb);
  if (t) time_end(st, 0.0);
}
void launch_acoustic_ring(const Args& a, double dts, int it, int ns, hipStream_t st) {
  const Geo& g = a.g;
  const unsigned ncol = unsigned(2 * g.nx + 2 * (g.ny - 2));
  const bool t = timing("acoustic_ring");
  if (t) time_begin(st);
  if (g.nz + 1 <= 64) hipLaunchKernelGGL(k_acoustic_ring<64>, dim3(ncol), dim3(64), 0, st, a, dts, it, ns);
  else if (g.nz + 1 <= 128) hipLaunchKernelGGL(k_acoustic_ring<128>, dim3(ncol), dim3(128), 0, st, a, dts, it, ns);
  else hipLaunchKernelGGL(k_acoustic_ring<256>, dim3(ncol), dim3(256), 0, st, a, dts, it, ns);
  if (t) time_end(st, 0.0);
}
void launch_bdy_zero_grad(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  const int maxpos = std::max(g.nx, g.ny) + 1;
  hipLaunchKernelGGL(k_bdy_zero_grad, dim3(unsigned((maxpos + 63) / 64), unsigned(g.nz + 1), 4), dim3(64, 1, 1), 0, st, a);
}
void launch_bdy_flow_dep(const Args& a, const Scalars& s, const int* sp, int n, hipStream_t st) {
  if (n <= 0) return;
  SpList l{};
  l.n = n;
  for (int q = 0; q < n; ++q) l.sp[q] = sp[q];
  const Geo& g = a.g;
  const int maxpos = std::max(g.nx, g.ny) + 1;
  hipLaunchKernelGGL(k_bdy_flow_dep, dim3(unsigned((maxpos + 63) / 64), unsigned(g.nz), unsigned(4 * n)), dim3(64, 1, 1), 0, st,
                     a, s, l);
}
void launch_bdy_time(const Args& a, hipStream_t st) {
  hipLaunchKernelGGL(k_bdy_time, dim3(1, 1, 1), dim3(64, 1, 1), 0, st, a.f.bdy_time, a.q.dt);
}
void launch_copy_scalars(const Args& a, const Scalars& s, int n, int64_t n3, hipStream_t st) {
  if (n <= 0) return;
  const int64_t off = int64_t(kLpad) + int64_t(a.g.hy) * a.g.sj;
  hipLaunchKernelGGL(k_copy_scalars, dim3(unsigned((n3 + 255) / 256), unsigned(n), 1), dim3(256, 1, 1), 0, st, s, n3, off);
}
void launch_microphysics(const Args& a, const Scalars& s, hipStream_t st) {
  const Geo& g = a.g;
  const double dt = a.q.dt;
  // moist_physics_prep over every mass point (q, t, p, al, ph -> the prep fields), Kessler over the columns
  // outside the spec zone, moist_physics_finish over every mass point
  launch3a("mp_prep", 12, k_mp_prep, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a, s);
  const int sz = g.spec() ? 1 : 0, i0 = sz, i1 = g.nx - sz, j0 = sz, j1 = g.ny - sz;
  if (i1 > i0 && j1 > j0)
    hipLaunchKernelGGL(k_kessler, dim3(unsigned((i1 - i0 + 63) / 64), unsigned(j1 - j0), 1), dim3(64, 1, 1), 0, st, a, s,
                       dt, i0, i1, j0, j1);
  launch3a("mp_finish", 9, k_mp_finish, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a, s, dt);
}
void launch_p_rho_phi(const Args& a, hipStream_t st) {
  const Geo& g = a.g;
  launch3("p_rho_phi", k_p_rho_phi, Range{0, g.nx, 0, g.ny, 0, g.nz}, st, a);
}
void launch_p_rho_phi_range(const Args& a, int i0, int i1, int j0, int j1, hipStream_t st) {
  if (i1 > i0 && j1 > j0) launch3("p_rho_phi", k_p_rho_phi, Range{i0, i1, j0, j1, 0, a.g.nz}, st, a);
}
void launch_finite(const Args& a, int* flags, hipStream_t st) {
  const Geo& g = a.g;
  launch3("finite", k_finite, Range{0, g.nx + 1, 0, g.nyv, 0, g.nz + 1}, st, a, flags);
}

}  // namespace MI_NS
