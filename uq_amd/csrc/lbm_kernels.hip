// lbm_kernels.hip — gfx950 kernels of the D2Q13 thermal FD-LBM.
//
// One step of the reference's RKMETHOD = 1 stepper (IMEX-SSP2(3,2,2),
// reference RK3, cpuLBM.cpp:1738-2103) is two tile launches, each covering
// every tile of every instance of the group in one grid-stride (persistent)
// loop, plus the wall closure when the group has walls:
//
//   [k_s12_list  f2 of the cells the wall closure's stencils read]
//   [k_alpha     wall cells: closure factor from the fluid neighbours' face
//                fluxes of f2 (reference wallFluxes, cpuLBM.cpp:1502-1590)]
//   k_stageA     stages 1-3 fused: f2 of tile + halo recomputed in LDS per
//                direction, face fluxes, moments of f - dF(f2), stage-3
//                equilibrium -> f3; G (stage-2 half of the final
//                combination) -> the other state buffer
//   [k_alpha     same for f3]
//   k_stageB     face fluxes of f3, stage-3 relaxation -> new state
//
// Face fluxes are staged one direction at a time: the tile's cells plus a
// 3-cell halo cross (rows for x fluxes, columns for y fluxes) of one f and
// one h plane, so LDS holds 2 x 22 x 22 values for a 16 x 16 tile whatever
// the precision. Halo positions outside the grid are resolved with the
// reference's index() rules (periodic wrap, mirror with direction remap,
// zeroth-order clamp; cpuLBM.cpp:1033-1169) when the tile is set up.
#include <climits>
#include <type_traits>

#include "lbm_device.hpp"
#include "lbm_launch.hpp"

namespace uqlbm {

constexpr int kTX = 16, kTY = 16, kH = 3;
constexpr int kNT = kTX * kTY;
constexpr int kRX = kTX + 2 * kH, kRY = kTY + 2 * kH;

// region meta bits
enum : unsigned {
  kMetaType = 3u,
  kMetaMirEW = 4u,   // x position reached through an east/west mirror
  kMetaMirNS = 8u,   // y position reached through a north/south mirror
  kMetaClampX = 16u, // x position clamped (zeroth-order extrapolation edge)
  kMetaClampY = 32u,
  kMetaUsed = 64u,
};

template <typename T>
__device__ __forceinline__ T* plane(void* base, long long plane_elems, int k) {
  return reinterpret_cast<T*>(base) + plane_elems * (long long)k;
}

// Reference index() along x/y (cpuLBM.cpp:1033-1113) for one position,
// followed by a clamp into the grid for positions no stencil reaches.
__device__ __forceinline__ int map_axis(int X, int n, int periodic, int mir_lo, int mir_hi,
                                        unsigned& mir, unsigned& clamped) {
  mir = 0;
  clamped = 0;
  if (X < 0) {
    if (periodic) X += n;
    else if (mir_lo) { X = -X - 1; mir = 1; }
    else { X = 0; clamped = 1; }
  } else if (X > n - 1) {
    if (periodic) X -= n;
    else if (mir_hi) { X = 2 * n - X - 1; mir = 1; }
    else { X = n - 1; clamped = 1; }
  }
  return X < 0 ? 0 : (X > n - 1 ? n - 1 : X);
}

// Grid-stride tile loop with an XCD-aware order: gfx950 hands consecutive
// workgroups to its 8 XCDs in turn, so block b is renumbered so that each
// XCD walks a contiguous run of tiles (neighbouring tiles share halos and
// the XCD's L2).
__device__ __forceinline__ int first_tile() {
  const int G = gridDim.x, b = blockIdx.x;
  if (G % 8 != 0) return b;
  return (b % 8) * (G / 8) + b / 8;
}

// ---------------------------------------------------------------------------
// Tile region: physical offsets and meta of the tile plus halo cross.
// ---------------------------------------------------------------------------
struct Region {
  int off[kRY][kRX];
  unsigned char meta[kRY][kRX];
};

__device__ void setup_region(Region& R, const InstDev& in, int x0, int y0) {
  const int px_ = in.flags & kPeriodicX, py_ = in.flags & kPeriodicY;
  const int mw = in.flags & kMirrorW, me = in.flags & kMirrorE;
  const int ms = in.flags & kMirrorS, mn = in.flags & kMirrorN;
  for (int r = threadIdx.x; r < kRX * kRY; r += kNT) {
    const int ly = r / kRX - kH, lx = r % kRX - kH;
    const bool inx = lx >= 0 && lx < kTX, iny = ly >= 0 && ly < kTY;
    if (!inx && !iny) {  // corner: never read
      R.off[ly + kH][lx + kH] = 0;
      R.meta[ly + kH][lx + kH] = 0;
      continue;
    }
    unsigned mx, cx, my, cy;
    const int X = map_axis(x0 + lx, in.nx, px_, mw, me, mx, cx);
    const int Y = map_axis(y0 + ly, in.ny, py_, ms, mn, my, cy);
    const long long o = (long long)Y * in.pitch + X;
    R.off[ly + kH][lx + kH] = (int)o;
    unsigned m = in.ctype[o] & kMetaType;
    m |= mx ? kMetaMirEW : 0u;
    m |= my ? kMetaMirNS : 0u;
    m |= cx ? kMetaClampX : 0u;
    m |= cy ? kMetaClampY : 0u;
    R.meta[ly + kH][lx + kH] = (unsigned char)(m | kMetaUsed);
  }
}

__device__ __forceinline__ int remap_dir(int i, unsigned m) {
  if (m & kMetaMirEW) i = mirror_ew(i);
  if (m & kMetaMirNS) i = mirror_ns(i);
  return i;
}

// Staging buffers and face fluxes of one direction.
template <typename TC>
struct Stage {
  TC f[kRY][kRX];
  TC h[kRY][kRX];
  TC fxf[kTY][kTX + 1], fxh[kTY][kTX + 1];
  TC fyf[kTY + 1][kTX], fyh[kTY + 1][kTX];
};

// Copy plane I (mirror-remapped per position) of `src` into the staging
// buffers for the own rect and the halo bands direction I needs.
template <int I, typename TS, typename TC>
__device__ __forceinline__ void stage_raw(Stage<TC>& S, const Region& R, const InstDev& in, void* src) {
  constexpr int EX = lat_x(I), EY = lat_y(I);
  constexpr int nown = kNT, nxb = EX ? 2 * kH * kTY : 0, nyb = EY ? 2 * kH * kTX : 0;
  for (int r = threadIdx.x; r < nown + nxb + nyb; r += kNT) {
    int lx, ly;
    if (r < nown) {
      lx = r % kTX;
      ly = r / kTX;
    } else if (r < nown + nxb) {
      const int k = r - nown, side = k / (kH * kTY), kk = k % (kH * kTY);
      ly = kk / kH;
      lx = side ? kTX + kk % kH : -kH + kk % kH;
    } else {
      const int k = r - nown - nxb, side = k / (kH * kTX), kk = k % (kH * kTX);
      lx = kk % kTX;
      ly = side ? kTY + kk / kTX : -kH + kk / kTX;
    }
    const unsigned m = R.meta[ly + kH][lx + kH];
    const int d = remap_dir(I, m);
    const long long o = R.off[ly + kH][lx + kH];
    S.f[ly + kH][lx + kH] = TC(plane<TS>(src, in.plane, d)[o]);
    S.h[ly + kH][lx + kH] = TC(plane<TS>(src, in.plane, kQ + d)[o]);
  }
}

// Reference stencil solid handling along one line (cpuLBM.cpp:1242-1322):
// the first solid/fluid transition's solid point becomes 2*V(m1) - V(m2) of
// the two nearest fluid points, and the point beyond it the same value.
// The addressing equivalence with the reference (positions inside the
// stencil, unmirrored direction) is checked on the host per geometry.
template <int L, typename TC>
__device__ __forceinline__ void extrapolate(const int* sld, const TC* V, TC* S) {
  int diff = 0, jj = 0;
#pragma unroll
  for (int j = 0; j < L - 1; j++) {
    const int d = sld[j] - sld[j + 1];
    if (d != 0) {
      diff = d;
      jj = d > 0 ? j : j + 1;
      break;
    }
  }
  int k1 = jj + diff, k2 = jj + 2 * diff;
  k1 = k1 < 0 ? 0 : (k1 > L - 1 ? L - 1 : k1);
  k2 = k2 < 0 ? 0 : (k2 > L - 1 ? L - 1 : k2);
  S[jj] = TC(2) * V[k1] - V[k2];
  const int nx = jj - diff;
  if (nx >= 0 && nx < L) S[nx] = S[jj];
}

// Face flux (f and h) out of the owner at logical (lx, ly) along axis AX for
// direction I: the wall closure at solid owners, else WENO5/NND on the
// staged line.
template <int I, int AX, typename TC>
__device__ __forceinline__ void face_flux(const Stage<TC>& S, const Region& R, const InstDev& in,
                                          bool solids, int lx, int ly, TC& vf, TC& vh) {
  constexpr int E = AX == 0 ? lat_x(I) : lat_y(I);
  constexpr int SG = sgn_i(E);
  const TC v = TC(double(E > 0 ? E : -E) * in.c);
  const unsigned mo = R.meta[ly + kH][lx + kH];
  if ((mo & kMetaType) == 2) {
    const TC a = TC(in.alpha[R.off[ly + kH][lx + kH]]);
    vf = a * v * S.f[ly + kH][lx + kH];
    vh = a * v * S.h[ly + kH][lx + kH];
    return;
  }
  if (in.flux_method == 1) {
    TC Vf[5], Vh[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const int px = AX == 0 ? lx + (k - 2) * SG : lx;
      const int py = AX == 0 ? ly : ly + (k - 2) * SG;
      Vf[k] = S.f[py + kH][px + kH];
      Vh[k] = S.h[py + kH][px + kH];
    }
    if (solids) {
      int sld[5], any = 0;
#pragma unroll
      for (int k = 0; k < 5; k++) {
        const int px = AX == 0 ? lx + (k - 2) * SG : lx;
        const int py = AX == 0 ? ly : ly + (k - 2) * SG;
        sld[k] = (R.meta[py + kH][px + kH] & kMetaType) == 2;
        any |= sld[k];
      }
      if (any) {
        TC Sf[5], Sh[5];
#pragma unroll
        for (int k = 0; k < 5; k++) { Sf[k] = Vf[k]; Sh[k] = Vh[k]; }
        extrapolate<5>(sld, Vf, Sf);
        extrapolate<5>(sld, Vh, Sh);
        vf = weno5(v, Sf[0], Sf[1], Sf[2], Sf[3], Sf[4]);
        vh = weno5(v, Sh[0], Sh[1], Sh[2], Sh[3], Sh[4]);
        return;
      }
    }
    vf = weno5(v, Vf[0], Vf[1], Vf[2], Vf[3], Vf[4]);
    vh = weno5(v, Vh[0], Vh[1], Vh[2], Vh[3], Vh[4]);
  } else {
    TC Vf[3], Vh[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int px = AX == 0 ? lx + (k - 1) * SG : lx;
      const int py = AX == 0 ? ly : ly + (k - 1) * SG;
      Vf[k] = S.f[py + kH][px + kH];
      Vh[k] = S.h[py + kH][px + kH];
    }
    if (solids) {
      int sld[3], any = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int px = AX == 0 ? lx + (k - 1) * SG : lx;
        const int py = AX == 0 ? ly : ly + (k - 1) * SG;
        sld[k] = (R.meta[py + kH][px + kH] & kMetaType) == 2;
        any |= sld[k];
      }
      if (any) {
        TC Sf[3], Sh[3];
#pragma unroll
        for (int k = 0; k < 3; k++) { Sf[k] = Vf[k]; Sh[k] = Vh[k]; }
        extrapolate<3>(sld, Vf, Sf);
        extrapolate<3>(sld, Vh, Sh);
        vf = nnd(v, Sf[0], Sf[1], Sf[2]);
        vh = nnd(v, Sh[0], Sh[1], Sh[2]);
        return;
      }
    }
    vf = nnd(v, Vf[0], Vf[1], Vf[2]);
    vh = nnd(v, Vh[0], Vh[1], Vh[2]);
  }
}

// All face fluxes direction I needs for the tile's cells: each cell's own
// out-face plus one upwind owner per row (x) / column (y).
template <int I, typename TC>
__device__ __forceinline__ void tile_faces(Stage<TC>& S, const Region& R, const InstDev& in, bool solids) {
  constexpr int EX = lat_x(I), EY = lat_y(I);
  constexpr int SX = sgn_i(EX), SY = sgn_i(EY);
  constexpr int nfx = EX ? (kTX + 1) * kTY : 0;
  constexpr int nfy = EY ? (kTY + 1) * kTX : 0;
  for (int k = threadIdx.x; k < nfx + nfy; k += kNT) {
    if (k < nfx) {
      const int row = k / (kTX + 1), col = k % (kTX + 1);
      const int lo = SX > 0 ? -1 : 0;
      TC vf, vh;
      face_flux<I, 0>(S, R, in, solids, lo + col, row, vf, vh);
      S.fxf[row][col] = vf;
      S.fxh[row][col] = vh;
    } else {
      const int kk = k - nfx;
      const int row = kk / kTX, col = kk % kTX;
      const int lo = SY > 0 ? -1 : 0;
      TC vf, vh;
      face_flux<I, 1>(S, R, in, solids, col, lo + row, vf, vh);
      S.fyf[row][col] = vf;
      S.fyh[row][col] = vh;
    }
  }
}

// Net flux of direction I at the own cell (reference cFLUX,
// cpuLBM.cpp:1436-1470): (dt/dx)(Fx_out - Fx_in) + (dt/dy)(Fy_out - Fy_in),
// grouped as out - in. An upwind owner across a clamped edge is the cell
// itself.
template <int I, typename TC>
__device__ __forceinline__ void own_net_flux(const Stage<TC>& S, const Region& R, const InstDev& in,
                                             int lx, int ly, TC& dF, TC& dH) {
  constexpr int EX = lat_x(I), EY = lat_y(I);
  constexpr int SX = sgn_i(EX), SY = sgn_i(EY);
  const TC rdx = TC(in.rdx), rdy = TC(in.rdy);
  TC oxf = 0, oxh = 0, ixf = 0, ixh = 0, oyf = 0, oyh = 0, iyf = 0, iyh = 0;
  if constexpr (EX != 0) {
    const int lo = SX > 0 ? -1 : 0;
    oxf = S.fxf[ly][lx - lo];
    oxh = S.fxh[ly][lx - lo];
    const int ux = lx - SX;
    if (R.meta[ly + kH][ux + kH] & kMetaClampX) {
      ixf = oxf;
      ixh = oxh;
    } else {
      ixf = S.fxf[ly][ux - lo];
      ixh = S.fxh[ly][ux - lo];
    }
  }
  if constexpr (EY != 0) {
    const int lo = SY > 0 ? -1 : 0;
    oyf = S.fyf[ly - lo][lx];
    oyh = S.fyh[ly - lo][lx];
    const int uy = ly - SY;
    if (R.meta[uy + kH][lx + kH] & kMetaClampY) {
      iyf = oyf;
      iyh = oyh;
    } else {
      iyf = S.fyf[uy - lo][lx];
      iyh = S.fyh[uy - lo][lx];
    }
  }
  if constexpr (EX != 0 && EY != 0) {
    dF = (rdx * oxf + rdy * oyf) - (rdx * ixf + rdy * iyf);
    dH = (rdx * oxh + rdy * oyh) - (rdx * ixh + rdy * iyh);
  } else if constexpr (EX != 0) {
    dF = rdx * oxf - rdx * ixf;
    dH = rdx * oxh - rdx * ixh;
  } else {
    dF = rdy * oyf - rdy * iyf;
    dH = rdy * oyh - rdy * iyh;
  }
}

template <int... Is>
struct Seq {};
template <int N, int... Is>
struct MakeSeq : MakeSeq<N - 1, N - 1, Is...> {};
template <int... Is>
struct MakeSeq<0, Is...> {
  using type = Seq<Is...>;
};

// `step` counts the steps of the current launch sequence from 1; the global
// step index of an instance is in.step_base + step.
__device__ __forceinline__ bool frozen_by_failure(const InstDev& in, int step) {
  return *in.fail < in.step_base + step;
}

__device__ __forceinline__ void flag_failure(const InstDev& in, int step, double rho, double sumfeq) {
  if (!(fabs(sumfeq - rho) <= 1e-6) || !isfinite(rho)) {
    // reference check: |sum feq - rho| > 1e-6 (cpuLBM.cpp:825-843); a
    // non-finite density also fails here (the reference misses NaN).
    atomicMin(in.fail, in.step_base + step);
  }
}

// ---------------------------------------------------------------------------
// RK3 stages 1 and 2 (cpuLBM.cpp:1819-1924) are pointwise. Their per-cell
// coefficients live in a record; positions reached through a mirror store the
// mirrored velocity, since eq_I(u -> -u) == eq_{mirror_ew(I)}(u) and likewise
// for v, so one direction I serves every halo position.
// ---------------------------------------------------------------------------
template <typename TC>
struct Rec12 {
  EqIn<TC> q;
  TC cU, cV;          // c*U, c*V: e_i . u = lat_x(i)*cU + lat_y(i)*cV
  TC af, ah, ahf, g;  // dt/(2 tf), dt/(2 th), dt/(2 thf), 1/thf
};

template <typename TC>
__device__ __forceinline__ Rec12<TC> make_rec12(const InstDev& in, Macro m, bool flip_x, bool flip_y) {
  if (flip_x) m.ux = -m.ux;
  if (flip_y) m.uy = -m.uy;
  Rec12<TC> r;
  r.q = eq_inputs<TC>(in, m);
  r.cU = TC(in.c) * TC(m.ux);
  r.cV = TC(in.c) * TC(m.uy);
  const TC dt = TC(in.dt);
  r.af = dt / (TC(2) * TC(m.tf));
  r.ah = dt / (TC(2) * TC(m.th));
  r.ahf = dt / (TC(2) * TC(m.thf));
  r.g = TC(1) / TC(m.thf);
  return r;
}

// f2, h2 of direction I from the raw f, h and the cell record; also returns
// the stage-1 equilibria (feq2 = feq1, reference cpuLBM.cpp:1874-1875).
template <int I, typename TC>
__device__ __forceinline__ void s12(const Rec12<TC>& r, TC hdt, TC fv, TC hv, TC& fe, TC& he, TC& f2, TC& h2) {
  eq_dir<I>(r.q, fe, he);
  const TC f1 = (fv + r.af * fe) / (TC(1) + r.af);
  const TC eu = TC(lat_x(I)) * r.cU + TC(lat_y(I)) * r.cV;
  const TC h1 = (hv - r.ahf * eu * (fe - f1) + r.ah * he) / (TC(1) + r.ah);
  f2 = (fv - r.af * (fe - f1) + r.af * fe) / (TC(1) + r.af);
  h2 = (hv - hdt * ((eu * r.g) * (fe - f2) - (eu * r.g) * (fe - f1)) - r.ah * (he - h1) + r.ah * he) /
       (TC(1) + r.ah);
}

template <typename TS>
__device__ __forceinline__ void load26(const TS* src, long long P, long long o, double* f, double* h) {
#pragma unroll
  for (int i = 0; i < kQ; i++) {
    f[i] = double(src[P * i + o]);
    h[i] = double(src[P * (kQ + i) + o]);
  }
}

__device__ __forceinline__ void moments26(const InstDev& in, const double* f, const double* h, Macro& m) {
  double rho = 0, rux = 0, ruy = 0, sh = 0;
#pragma unroll
  for (int i = 0; i < kQ; i++) {
    rho += f[i];
    rux += f[i] * (lat_x(i) * in.c);
    ruy += f[i] * (lat_y(i) * in.c);
    sh += h[i];
  }
  finish_macro(in, rho, rux, ruy, sh, m);
}

// Halo positions of a tile: the x band (H columns either side, tile rows)
// then the y band (H rows either side, tile columns).
constexpr int kNXB = 2 * kH * kTY, kNYB = 2 * kH * kTX, kNHalo = kNXB + kNYB;

__device__ __forceinline__ void halo_pos(int k, int& lx, int& ly) {
  if (k < kNXB) {
    const int side = k / (kH * kTY), kk = k % (kH * kTY);
    ly = kk / kH;
    lx = side ? kTX + kk % kH : -kH + kk % kH;
  } else {
    const int k2 = k - kNXB, side = k2 / (kH * kTX), kk = k2 % (kH * kTX);
    lx = kk % kTX;
    ly = side ? kTY + kk / kTX : -kH + kk / kTX;
  }
}

// ---------------------------------------------------------------------------
// k_s12_list: f2 for a list of cells (the stencil cells of the wall closure).
// ---------------------------------------------------------------------------
template <typename TS, typename TC, int... Is>
__device__ __forceinline__ void s12_cell(const InstDev& in, const TS* src, TS* dst, long long o, Seq<Is...>) {
  const long long P = in.plane;
  double f[kQ], h[kQ];
  load26(src, P, o, f, h);
  Macro m;
  moments26(in, f, h, m);
  const Rec12<TC> r = make_rec12<TC>(in, m, false, false);
  const TC hdt = TC(in.dt) / TC(2);
  auto one = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    TC fe, he, f2, h2;
    s12<I>(r, hdt, TC(f[I]), TC(h[I]), fe, he, f2, h2);
    dst[P * I + o] = TS(f2);
    dst[P * (kQ + I) + o] = TS(h2);
  };
  (one(std::integral_constant<int, Is>{}), ...);
}

template <typename TS, typename TC>
__global__ __launch_bounds__(256) void k_s12_list(const InstDev* insts, const WallRef* cells, int ncells, int par,
                                                   int step) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncells) return;
  const WallRef cr = cells[k];
  const InstDev& in = insts[cr.inst];
  if (frozen_by_failure(in, step)) return;
  const long long o = (long long)cr.y * in.pitch + cr.x;
  const TS* src = reinterpret_cast<const TS*>(in.state[in.parity0 ^ par]);
  TS* dst = reinterpret_cast<TS*>(in.f2);
  if (in.ctype[o] > 0) {
#pragma unroll
    for (int q = 0; q < kPlanes; q++) dst[in.plane * q + o] = src[in.plane * q + o];
    return;
  }
  s12_cell<TS, TC>(in, src, dst, o, MakeSeq<kQ>::type{});
}

// ---------------------------------------------------------------------------
// k_stageA: RK3 stages 1-3 fused (cpuLBM.cpp:1819-1993) plus the stage-2
// half of the final combination (:2022, :2039).
//
// Per tile: stage-1/2 records for the halo cross go to LDS; per direction
// the f2/h2 plane of tile + halo is computed into the staging buffer, face
// fluxes follow, and each cell keeps two values (P = f - dF2 + af2 (feq2 -
// f2), Q likewise for h) while G = f - dF2/2 + dt/2 (feq2 - f2)/tf2 (and the
// h analogue) goes straight to the output state buffer. After the last
// direction the moments of f - dF(f2) give the stage-3 equilibrium and f3.
// ---------------------------------------------------------------------------
template <typename TC>
struct HaloStore {
  Rec12<TC> rec[kNHalo];
};

template <typename TS, typename TC, int... Is>
__device__ __forceinline__ void stageA_tile(Stage<TC>& S, HaloStore<TC>& HS, const Region& R, const InstDev& in,
                                            const TileRef& tr, int par, int step, Seq<Is...>) {
  const int lx = (int)threadIdx.x % kTX, ly = (int)threadIdx.x / kTX;
  const int x = tr.tx * kTX + lx, y = tr.ty * kTY + ly;
  const bool own_in = x < in.nx && y < in.ny;
  const unsigned mo = R.meta[ly + kH][lx + kH];
  const long long oo = R.off[ly + kH][lx + kH];   // mapped physical cell of the own position
  const int type_o = mo & kMetaType;
  const bool fluid = own_in && type_o == 0;
  const bool solids = tr.flags & kTileSolid;
  const long long P = in.plane;
  const TS* src = reinterpret_cast<const TS*>(in.state[in.parity0 ^ par]);
  TS* out = reinterpret_cast<TS*>(in.state[1 ^ in.parity0 ^ par]);
  TS* f3buf = reinterpret_cast<TS*>(in.f3);
  const TC dt = TC(in.dt), hdt = dt / TC(2);

  // --- phase 1: own record (registers) and halo records (LDS) -------------
  double fr[kQ], hr[kQ];
  load26(src, P, oo, fr, hr);
  Macro m1;
  moments26(in, fr, hr, m1);
  const Rec12<TC> ro = make_rec12<TC>(in, m1, mo & kMetaMirEW, mo & kMetaMirNS);
  for (int k = threadIdx.x; k < kNHalo; k += kNT) {
    int hx, hy;
    halo_pos(k, hx, hy);
    const unsigned mh = R.meta[hy + kH][hx + kH];
    if ((mh & kMetaType) != 0) continue;
    double f[kQ], h[kQ];
    load26(src, P, R.off[hy + kH][hx + kH], f, h);
    Macro m;
    moments26(in, f, h, m);
    HS.rec[k] = make_rec12<TC>(in, m, mh & kMetaMirEW, mh & kMetaMirNS);
  }
  __syncthreads();

  TC Pf[kQ], Qh[kQ];
  double rho3 = 0, rux3 = 0, ruy3 = 0, sh3 = 0;
  const TC tf2 = TC(m1.tf), th2 = TC(m1.th), thf2 = TC(m1.thf);
  const TC af2 = dt / (TC(2) * tf2), ah2 = dt / (TC(2) * th2);

  auto one_dir = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    constexpr int EX = lat_x(I), EY = lat_y(I);
    // own position: f2/h2 from the record (raw value of the mirrored direction
    // when the position lies outside the grid)
    TC fe = 0, he = 0, f2 = 0, h2 = 0;
    {
      const int d = remap_dir(I, mo);
      const TC fv = own_in ? TC(fr[I]) : TC(plane<TS>((void*)src, P, d)[oo]);
      const TC hv = own_in ? TC(hr[I]) : TC(plane<TS>((void*)src, P, kQ + d)[oo]);
      if (type_o == 0) {
        s12<I>(ro, hdt, fv, hv, fe, he, f2, h2);
      } else {
        f2 = fv;
        h2 = hv;
      }
    }
    if constexpr (I != 0) {
      S.f[ly + kH][lx + kH] = f2;
      S.h[ly + kH][lx + kH] = h2;
      constexpr int k0 = EX ? 0 : kNXB;
      constexpr int k1 = EY ? kNHalo : kNXB;
      for (int k = k0 + (int)threadIdx.x; k < k1; k += kNT) {
        int hx, hy;
        halo_pos(k, hx, hy);
        const unsigned mh = R.meta[hy + kH][hx + kH];
        const long long oh = R.off[hy + kH][hx + kH];
        const int d = remap_dir(I, mh);
        const TC fv = TC(plane<TS>((void*)src, P, d)[oh]);
        const TC hv = TC(plane<TS>((void*)src, P, kQ + d)[oh]);
        TC a = fv, b = hv;
        if ((mh & kMetaType) == 0) {
          TC fe_, he_;
          s12<I>(HS.rec[k], hdt, fv, hv, fe_, he_, a, b);
        }
        S.f[hy + kH][hx + kH] = a;
        S.h[hy + kH][hx + kH] = b;
      }
      __syncthreads();
      tile_faces<I, TC>(S, R, in, solids);
      __syncthreads();
    }
    if (fluid) {
      TC dF = 0, dH = 0;
      if constexpr (I != 0) own_net_flux<I, TC>(S, R, in, lx, ly, dF, dH);
      const TC fv = TC(fr[I]), hv = TC(hr[I]);
      const TC D = fe - f2;
      const TC eu2 = TC(lat_x(I)) * ro.cU + TC(lat_y(I)) * ro.cV;
      const TC A2 = (eu2 / thf2) * D;
      Pf[I] = (fv - dF) + af2 * D;
      Qh[I] = ((hv - dH) - hdt * A2) + ah2 * (he - h2);
      const TC gf = (fv - TC(0.5) * dF) + hdt * (D / tf2);
      const TC gh = ((hv - TC(0.5) * dH) + hdt * ((he - h2) / th2)) - hdt * A2;
      out[P * I + oo] = TS(gf);
      out[P * (kQ + I) + oo] = TS(gh);
      const double fd = fr[I], dd = double(dF);
      const double exc = lat_x(I) * in.c, eyc = lat_y(I) * in.c;
      rho3 += fd - dd;
      rux3 += fd * exc - dd * exc;
      ruy3 += fd * eyc - dd * eyc;
      sh3 += hr[I] - double(dH);
    } else if (own_in) {
      // frozen cell: f3 = f2 = f and the new state is f
      const TS a = src[P * I + oo], b = src[P * (kQ + I) + oo];
      out[P * I + oo] = a;
      out[P * (kQ + I) + oo] = b;
      f3buf[P * I + oo] = a;
      f3buf[P * (kQ + I) + oo] = b;
    }
  };
  (one_dir(std::integral_constant<int, Is>{}), ...);

  if (!fluid) return;
  Macro m3;
  finish_macro(in, rho3, rux3, ruy3, sh3, m3);
  const EqIn<TC> q3 = eq_inputs<TC>(in, m3);
  const TC tf3 = TC(m3.tf), th3 = TC(m3.th), thf3 = TC(m3.thf);
  const TC af3 = dt / (TC(2) * tf3), ah3 = dt / (TC(2) * th3);
  const TC cU3 = TC(in.c) * TC(m3.ux), cV3 = TC(in.c) * TC(m3.uy);
  double sumfe = 0;
  auto finish = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    TC fe, he;
    eq_dir<I>(q3, fe, he);
    sumfe += double(fe);
    const TC f3 = (Pf[I] + af3 * fe) / (TC(1) + af3);
    const TC eu3 = TC(lat_x(I)) * cU3 + TC(lat_y(I)) * cV3;
    const TC A3 = (eu3 / thf3) * (fe - f3);
    const TC h3 = ((Qh[I] - hdt * A3) + ah3 * he) / (TC(1) + ah3);
    f3buf[P * I + oo] = TS(f3);
    f3buf[P * (kQ + I) + oo] = TS(h3);
  };
  (finish(std::integral_constant<int, Is>{}), ...);
  flag_failure(in, step, m3.rho, sumfe);
}

template <typename TS, typename TC>
__global__ __launch_bounds__(kNT) void k_stageA(const InstDev* insts, const TileRef* tiles, int ntiles, int par,
                                                int step) {
  __shared__ Region R;
  __shared__ Stage<TC> S;
  __shared__ HaloStore<TC> HS;
  for (int t = first_tile(); t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const InstDev& in = insts[tr.inst];
    if (frozen_by_failure(in, step)) continue;
    __syncthreads();
    setup_region(R, in, tr.tx * kTX, tr.ty * kTY);
    __syncthreads();
    stageA_tile<TS, TC>(S, HS, R, in, tr, par, step, MakeSeq<kQ>::type{});
  }
}

// ---------------------------------------------------------------------------
// k_stageB: the rest of the final combination (cpuLBM.cpp:1995-2042): the
// stage-3 relaxation terms (equilibrium from the moments of f3, which equal
// the moments of f - dF(f2)) and -dF(f3)/2.
// ---------------------------------------------------------------------------
template <typename TS, typename TC, int... Is>
__device__ __forceinline__ void stageB_tile(Stage<TC>& S, const Region& R, const InstDev& in, const TileRef& tr,
                                            int par, Seq<Is...>) {
  const int lx = (int)threadIdx.x % kTX, ly = (int)threadIdx.x / kTX;
  const int x = tr.tx * kTX + lx, y = tr.ty * kTY + ly;
  const bool own_in = x < in.nx && y < in.ny;
  const long long oo = R.off[ly + kH][lx + kH];
  const bool fluid = own_in && (R.meta[ly + kH][lx + kH] & kMetaType) == 0;
  const bool solids = tr.flags & kTileSolid;
  const long long P = in.plane;
  const TS* f3b = reinterpret_cast<const TS*>(in.f3);
  TS* out = reinterpret_cast<TS*>(in.state[1 ^ in.parity0 ^ par]);
  const TC dt = TC(in.dt), hdt = dt / TC(2);
  Macro m3{};
  m3.rho = 1;
  m3.tf = m3.th = m3.thf = 1;
  if (fluid) {
    double f[kQ], h[kQ];
    load26(f3b, P, oo, f, h);
    moments26(in, f, h, m3);
  }
  const EqIn<TC> q3 = eq_inputs<TC>(in, m3);
  const TC tf3 = TC(m3.tf), th3 = TC(m3.th), thf3 = TC(m3.thf);
  const TC cU3 = TC(in.c) * TC(m3.ux), cV3 = TC(in.c) * TC(m3.uy);
  auto one_dir = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    TC dF = 0, dH = 0;
    if constexpr (I != 0) {
      stage_raw<I, TS, TC>(S, R, in, in.f3);
      __syncthreads();
      tile_faces<I, TC>(S, R, in, solids);
      __syncthreads();
      if (fluid) own_net_flux<I, TC>(S, R, in, lx, ly, dF, dH);
    }
    if (fluid) {
      TC fe, he;
      eq_dir<I>(q3, fe, he);
      const TC f3 = TC(f3b[P * I + oo]), h3 = TC(f3b[P * (kQ + I) + oo]);
      const TC eu3 = TC(lat_x(I)) * cU3 + TC(lat_y(I)) * cV3;
      const TC A3 = (eu3 / thf3) * (fe - f3);
      TS* pf = plane<TS>(out, P, I);
      TS* ph = plane<TS>(out, P, kQ + I);
      pf[oo] = TS((TC(pf[oo]) + hdt * ((fe - f3) / tf3)) - TC(0.5) * dF);
      ph[oo] = TS(((TC(ph[oo]) + hdt * ((he - h3) / th3)) - hdt * A3) - TC(0.5) * dH);
    }
  };
  (one_dir(std::integral_constant<int, Is>{}), ...);
}

template <typename TS, typename TC>
__global__ __launch_bounds__(kNT) void k_stageB(const InstDev* insts, const TileRef* tiles, int ntiles, int par,
                                                int step) {
  __shared__ Region R;
  __shared__ Stage<TC> S;
  for (int t = first_tile(); t < ntiles; t += gridDim.x) {
    const TileRef tr = tiles[t];
    const InstDev& in = insts[tr.inst];
    if (frozen_by_failure(in, step)) continue;
    __syncthreads();
    setup_region(R, in, tr.tx * kTX, tr.ty * kTY);
    __syncthreads();
    stageB_tile<TS, TC>(S, R, in, tr, par, MakeSeq<kQ>::type{});
  }
}

// ---------------------------------------------------------------------------
// k_alpha: wall closure factor of each wall cell with a non-solid axis
// neighbour, from the neighbours' face fluxes of `which` (0: f2, 1: f3),
// following the reference's physical stencil logic (cpuLBM.cpp:1171-1326,
// 1502-1590).
// ---------------------------------------------------------------------------
__constant__ int c_wall_dx[4] = {-1, 1, 0, 0};
__constant__ int c_wall_dy[4] = {0, 0, -1, 1};
__constant__ int c_wall_in[4][4] = {{1, 5, 8, 9}, {3, 6, 7, 11}, {2, 5, 6, 10}, {4, 7, 8, 12}};
__constant__ int c_wall_out[4][4] = {{3, 6, 7, 11}, {1, 5, 8, 9}, {4, 8, 7, 12}, {2, 6, 5, 10}};
__constant__ int c_lat_x[13] = {0, 1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0};
__constant__ int c_lat_y[13] = {0, 0, 1, 0, -1, 1, 1, -1, -1, 0, 2, 0, -2};
__constant__ int c_mir_ns[13] = {0, 1, 4, 3, 2, 8, 7, 6, 5, 9, 12, 11, 10};
__constant__ int c_mir_ew[13] = {0, 3, 2, 1, 4, 6, 5, 8, 7, 11, 10, 9, 12};

__device__ __forceinline__ int ref_index(const InstDev& in, int i, int pm, int d, int& mir) {
  unsigned m, cl;
  int r;
  if (d == 0)
    r = map_axis(i + pm, in.nx, in.flags & kPeriodicX, in.flags & kMirrorW, in.flags & kMirrorE, m, cl);
  else
    r = map_axis(i + pm, in.ny, in.flags & kPeriodicY, in.flags & kMirrorS, in.flags & kMirrorN, m, cl);
  mir = (int)m;
  return r;
}

template <typename TS, typename TC>
__device__ TC ref_face_flux_f(const InstDev& in, const TS* src, int x, int y, int i, int axis) {
  const int e = axis == 0 ? c_lat_x[i] : c_lat_y[i];
  const int s = e > 0 ? 1 : (e < 0 ? -1 : 0);
  if (s == 0) return TC(0);
  const int L = in.flux_method == 1 ? 5 : 3;
  const int low = (L - 1) / 2;
  const int start = (axis == 0 ? x : y) - s * low;
  TC V[5];
  int sld[5], any = 0;
  for (int j = 0; j < L; j++) {
    int mir;
    const int p = ref_index(in, start, s * j, axis, mir);
    const int d = mir ? (axis == 0 ? c_mir_ew[i] : c_mir_ns[i]) : i;
    const long long o = axis == 0 ? (long long)y * in.pitch + p : (long long)p * in.pitch + x;
    V[j] = TC(src[in.plane * d + o]);
    sld[j] = in.ctype[o] == 2;
    any |= sld[j];
  }
  if (any) {
    int diff = 0, jj = 0, pos = 0, mir;
    for (int j = 0; j < L - 1; j++) {
      diff = sld[j] - sld[j + 1];
      if (diff > 0) { jj = j; pos = ref_index(in, start, s * j, axis, mir); break; }
      if (diff < 0) { jj = j + 1; pos = ref_index(in, start, s * (j + 1), axis, mir); break; }
    }
    const int m1 = ref_index(in, pos, s * diff, axis, mir);
    const int m2 = ref_index(in, pos, s * 2 * diff, axis, mir);
    const long long o1 = axis == 0 ? (long long)y * in.pitch + m1 : (long long)m1 * in.pitch + x;
    const long long o2 = axis == 0 ? (long long)y * in.pitch + m2 : (long long)m2 * in.pitch + x;
    V[jj] = TC(2) * TC(src[in.plane * i + o1]) - TC(src[in.plane * i + o2]);
    const int nx = jj - diff;
    if (nx >= 0 && nx < L) V[nx] = V[jj];
  }
  const TC v = TC(double(e > 0 ? e : -e) * in.c);
  if (L == 5) return weno5(v, V[0], V[1], V[2], V[3], V[4]);
  return nnd(v, V[0], V[1], V[2]);
}

template <typename TS, typename TC>
__global__ void k_alpha(const InstDev* insts, const WallRef* walls, int nwalls, int which, int step) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwalls) return;
  const WallRef wr = walls[w];
  const InstDev& in = insts[wr.inst];
  if (frozen_by_failure(in, step)) return;
  const TS* src = reinterpret_cast<const TS*>(which ? in.f3 : in.f2);
  const long long ow = (long long)wr.y * in.pitch + wr.x;
  double in_f = 0, out_f = 0;
  for (int j = 0; j < 4; j++) {
    int mx, my;
    const int xn = ref_index(in, wr.x, c_wall_dx[j], 0, my);
    const int yn = ref_index(in, wr.y, c_wall_dy[j], 1, mx);
    const long long on = (long long)yn * in.pitch + xn;
    if (in.ctype[on] >= 2) continue;
    const int axis = j < 2 ? 0 : 1;
    const double dd = axis == 0 ? in.dx : in.dy;
    for (int k = 0; k < 4; k++) {
      const int i = c_wall_in[j][k], io = c_wall_out[j][k];
      const double F = double(ref_face_flux_f<TS, TC>(in, src, xn, yn, i, axis));
      in_f += F / dd;
      const double vv = fabs(double(axis == 0 ? c_lat_x[io] : c_lat_y[io]) * in.c);
      const double fw = double(src[in.plane * io + ow]);
      out_f += fw * vv / dd;
    }
  }
  in.alpha[ow] = in_f / out_f;
}

// ---------------------------------------------------------------------------
// Initialisation (reference initialiseFunctions, cpuLBM.cpp:846-875) and
// macroscopic output (macroPropShort, cpuLBM.cpp:877-914).
// ---------------------------------------------------------------------------
template <typename TS, int... Is>
__device__ __forceinline__ void init_cell(const InstDev& in, TS* dst, long long o, const double* pr,
                                          Seq<Is...>) {
  Macro m;
  m.rho = pr[0];
  m.ux = pr[1];
  m.uy = pr[2];
  m.T = pr[3];
  const EqIn<double> q = eq_inputs<double>(in, m);
  auto one = [&](auto Ic) {
    constexpr int I = decltype(Ic)::value;
    double fe, he;
    eq_dir<I>(q, fe, he);
    dst[in.plane * I + o] = TS(fe);
    dst[in.plane * (kQ + I) + o] = TS(he);
  };
  (one(std::integral_constant<int, Is>{}), ...);
}

// props: per instance, rows of {rho, ux, uy, T}; ids: [ny][pitch]
template <typename TS>
__global__ void k_init(const InstDev* insts, int inst, const int* ids, const double* props) {
  const InstDev& in = insts[inst];
  const long long n = (long long)in.ny * in.pitch;
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n;
       o += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(o % in.pitch);
    if (x >= in.nx) continue;
    TS* dst = reinterpret_cast<TS*>(in.state[in.parity0]);
    init_cell<TS>(in, dst, o, props + 4 * ids[o], MakeSeq<kQ>::type{});
  }
}

// out: 5 planes of double [ny][pitch]: rho, ux, uy, p, T
template <typename TS>
__global__ void k_macro(const InstDev* insts, int inst, int par, double* outp) {
  const InstDev& in = insts[inst];
  const long long n = (long long)in.ny * in.pitch;
  const TS* src = reinterpret_cast<const TS*>(in.state[in.parity0 ^ par]);
  for (long long o = blockIdx.x * (long long)blockDim.x + threadIdx.x; o < n;
       o += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(o % in.pitch);
    if (x >= in.nx) continue;
    double rho = 0, rux = 0, ruy = 0, sh = 0;
#pragma unroll
    for (int i = 0; i < kQ; i++) {
      const double fi = double(src[in.plane * i + o]);
      rho += fi;
      rux += fi * (lat_x(i) * in.c);
      ruy += fi * (lat_y(i) * in.c);
      sh += double(src[in.plane * (kQ + i) + o]);
    }
    Macro m;
    finish_macro(in, rho, rux, ruy, sh, m);
    outp[o] = m.rho;
    outp[n + o] = m.ux;
    outp[2 * n + o] = m.uy;
    outp[3 * n + o] = m.p;
    outp[4 * n + o] = m.T;
  }
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------
int tile_w() { return kTX; }
int tile_h() { return kTY; }
int tile_halo() { return kH; }

template <typename TS, typename TC>
static int grid_for(int ntiles, const void* kernel) {
  int dev = 0, cus = 256, per = 4;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kNT, 0) != hipSuccess || per < 1) per = 1;
  long long g = (long long)cus * per;
  if (g > ntiles) g = ntiles;
  if (g >= 8) g -= g % 8;
  return (int)(g < 1 ? 1 : g);
}

template <typename TS, typename TC>
void launch_rk3_step(const StepLaunch& L, hipStream_t s, int par, int step, KernelTimer* tm) {
  static int gA = 0, gB = 0;
  if (!gA) gA = grid_for<TS, TC>(1 << 30, (const void*)k_stageA<TS, TC>);
  if (!gB) gB = grid_for<TS, TC>(1 << 30, (const void*)k_stageB<TS, TC>);
  const int n = L.ntiles;
  if (L.nwalls) {
    if (tm) tm->begin(0, s);
    k_s12_list<TS, TC><<<(L.ncells + 255) / 256, 256, 0, s>>>(L.insts, L.cells, L.ncells, par, step);
    if (tm) tm->end(0, s);
    if (tm) tm->begin(1, s);
    k_alpha<TS, TC><<<(L.nwalls + 255) / 256, 256, 0, s>>>(L.insts, L.walls, L.nwalls, 0, step);
    if (tm) tm->end(1, s);
  }
  if (tm) tm->begin(2, s);
  k_stageA<TS, TC><<<n < gA ? n : gA, kNT, 0, s>>>(L.insts, L.tiles, n, par, step);
  if (tm) tm->end(2, s);
  if (L.nwalls) {
    if (tm) tm->begin(1, s);
    k_alpha<TS, TC><<<(L.nwalls + 255) / 256, 256, 0, s>>>(L.insts, L.walls, L.nwalls, 1, step);
    if (tm) tm->end(1, s);
  }
  if (tm) tm->begin(3, s);
  k_stageB<TS, TC><<<n < gB ? n : gB, kNT, 0, s>>>(L.insts, L.tiles, n, par, step);
  if (tm) tm->end(3, s);
}

template <typename TS>
void launch_init(const InstDev* insts, int inst, const int* ids, const double* props, long long cells,
                 hipStream_t s) {
  const int blocks = (int)((cells + 255) / 256 < 4096 ? (cells + 255) / 256 : 4096);
  k_init<TS><<<blocks, 256, 0, s>>>(insts, inst, ids, props);
}

template <typename TS>
void launch_macro(const InstDev* insts, int inst, int par, double* out, long long cells, hipStream_t s) {
  const int blocks = (int)((cells + 255) / 256 < 4096 ? (cells + 255) / 256 : 4096);
  k_macro<TS><<<blocks, 256, 0, s>>>(insts, inst, par, out);
}

template void launch_rk3_step<double, double>(const StepLaunch&, hipStream_t, int, int, KernelTimer*);
template void launch_rk3_step<float, float>(const StepLaunch&, hipStream_t, int, int, KernelTimer*);
template void launch_init<double>(const InstDev*, int, const int*, const double*, long long, hipStream_t);
template void launch_init<float>(const InstDev*, int, const int*, const double*, long long, hipStream_t);
template void launch_macro<double>(const InstDev*, int, int, double*, long long, hipStream_t);
template void launch_macro<float>(const InstDev*, int, int, double*, long long, hipStream_t);

}  // namespace uqlbm
