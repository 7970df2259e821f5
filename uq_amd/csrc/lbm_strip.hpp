// lbm_strip.hpp — strip flux kernels (k_strip) of the D2Q13 thermal FD-LBM,
// included by lbm_kernels.hip after the face reconstructions.
//
// The net fluxes of MODE 0, 1 and 2 without moment sums (every flux stage
// except the stage-3 flux of FP32 storage) for the instances whose cells all
// take the logical stencil addressing. One wavefront per work item: a strip
// of kStripW columns and up to kStripH rows of one instance, marched row by
// row, one direction after the other. Lane l holds column x0 - 3 + l (a
// 3-column halo on each side fills the 64 lanes); a lane's y stencil is a
// window of six rows in its registers and its x stencil comes from the
// neighbouring lanes by whole-wave DPP shifts, so a work item needs no LDS
// and no barrier. Each face is computed once: the x face left of a lane's
// column by that lane, the y face below a row when the window reaches it;
// a cell's net flux takes the face it shares with its right neighbour from
// that lane and the face below it from the previous row. The arithmetic of
// every face and net flux is that of the tile kernels (face_flux, flux_dir).
// The next row's loads are issued before the current row's faces, so their
// latency hides behind them.
// (Included inside namespace uqlbm, after the gptr global-pointer alias.)
#pragma once

// DPP controls of whole-wave shifts by one lane (GFX9 wave_shr:1 and
// wave_shl:1): lane l takes the value of lane l - 1, respectively l + 1;
// the lane shifted in at the end reads 0 (bound_ctrl: the zero comes from
// the instruction, so no register has to be cleared before each shift, as
// an "old" operand of 0 would need). Every shift runs with all 64 lanes
// active (never inside a divergent branch).
constexpr int kDppShr1 = 0x138;
constexpr int kDppShl1 = 0x130;

template <int CTRL>
__device__ __forceinline__ int lane_shift(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float lane_shift(float v) {
  return __int_as_float(lane_shift<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double lane_shift(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_shift<CTRL>(int(b)), hi = lane_shift<CTRL>(int(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// One staged position of a direction: f, h and (solid items) its cell
// type, kept as loaded: compared only where used, so that the wait for the
// load falls there and not right after it (which would also wait for the
// loads of the next row in flight).
template <typename TC>
struct SPos {
  TC f, h;
  int type;
  TC a;  // solid items: the wall closure factor of the position
};

// Per-lane constants of a work item.
template <typename TC>
struct StripLane {
  int X;           // mapped column of the lane's logical column
  int mx;          // the column is mirrored east/west
  int own;         // an own column of the item inside the instance
  int clampL;      // the logical column left of the lane is a clamped position
  int clampR;      // ... right of the lane
  TC rdx;          // dt/dx of the lane's column (own lanes)
  int xl;          // logical column
};

// Face flux of the line s0..s4 (WENO5; NND takes s1..s3), with the solid
// extrapolation of face_flux when a point of the line is solid.
template <typename TC, bool SOLID, bool W5>
__device__ __forceinline__ void strip_line(TC v, const TC (&Vf)[5], const TC (&Vh)[5], const int (&sld)[5], TC& vf,
                                           TC& vh) {
  if constexpr (W5) {
    TC F[5], H[5];
#pragma unroll
    for (int k = 0; k < 5; k++) { F[k] = Vf[k]; H[k] = Vh[k]; }
    if constexpr (SOLID) {
      if (sld[0] | sld[1] | sld[2] | sld[3] | sld[4]) {
        TC Ef[5], Eh[5];
#pragma unroll
        for (int k = 0; k < 5; k++) { Ef[k] = F[k]; Eh[k] = H[k]; }
        extrapolate<5>(sld, F, Ef);
        extrapolate<5>(sld, H, Eh);
#pragma unroll
        for (int k = 0; k < 5; k++) { F[k] = Ef[k]; H[k] = Eh[k]; }
      }
    }
    vf = weno5(v, F[0], F[1], F[2], F[3], F[4]);
    vh = weno5(v, H[0], H[1], H[2], H[3], H[4]);
  } else {
    TC F[3] = {Vf[1], Vf[2], Vf[3]}, H[3] = {Vh[1], Vh[2], Vh[3]};
    if constexpr (SOLID) {
      const int s3[3] = {sld[1], sld[2], sld[3]};
      if (s3[0] | s3[1] | s3[2]) {
        TC Ef[3] = {F[0], F[1], F[2]}, Eh[3] = {H[0], H[1], H[2]};
        extrapolate<3>(s3, F, Ef);
        extrapolate<3>(s3, H, Eh);
#pragma unroll
        for (int k = 0; k < 3; k++) { F[k] = Ef[k]; H[k] = Eh[k]; }
      }
    }
    vf = nnd(v, F[0], F[1], F[2]);
    vh = nnd(v, H[0], H[1], H[2]);
  }
}

// Wall closure face of a solid owner (face_flux): alpha |e| f, for FP32
// storage (deviations dev = f - ref staged) |e| (alpha dev + (alpha - 1) ref).
template <typename TS, typename TC>
__device__ __forceinline__ void strip_wall(TC a, TC v, TC wf, TC wh, TC rf, TC rh, TC& vf, TC& vh) {
#pragma clang fp contract(off)  // rounds as the tile kernels' wall faces
  if constexpr (sizeof(TS) == 4) {
    vf = v * (a * wf + (a - TC(1)) * rf);
    vh = v * (a * wh + (a - TC(1)) * rh);
  } else {
    (void)rf;
    (void)rh;
    vf = a * v * wf;
    vh = a * v * wh;
  }
}

// The instance fields the strip kernels read, copied once per work item:
// read through the instance table after the item's stores they would be
// reloaded (the stores might alias the table).
template <typename TS>
struct StripInst {
  gptr<const TS> src;
  gptr<TS> dst;
  gptr<const TS> dfl;  // MODE 1: the stored stage-2 net fluxes
  gptr<const uint8_t> ct;
  gptr<const double> alpha;
  gptr<const double> rdyp;  // dt/dy per row (rdy_stride 1) or the instance's one value (0)
  int rdy_stride;
  long long P;
  int pitch, ny, py, ms, mn;
  int inner;  // the item's rows and their 3-row halo lie inside the instance
  double c;
};

// One direction over the rows of a work item. SX, SY: the signs of the
// direction's lattice velocity in the march's row order; SOLID: the item's
// region holds a solid cell; W5: WENO5 faces (else NND). flip: the march
// takes the item's rows from the last to the first, as rows y0 .. y0 +
// nrows - 1 of the item turned upside down (row y of the march is row
// 2 y0 + nrows - 1 - y of the instance), called with SY the negated sign
// of the direction: each face then has the same upwind owner and the same
// stencil values in the same order as in the ascending march, so the two
// orders give the same values bit for bit.
template <typename TS, typename TC, int MODE, bool SOLID, bool W5, int SX, int SY>
__device__ __forceinline__ void strip_dir(const StripInst<TS>& in, const StripLane<TC>& C, TC rf, TC rh, int I,
                                          int y0, int nrows, int flip) {
  constexpr bool HX = SX != 0, HY = SY != 0;
  // the instance row of row y of the march
  const int yflip = 2 * y0 + nrows - 1;
  auto prow = [&](int y) { return flip ? yflip - y : y; };
  const int pitch = in.pitch, ny = in.ny;
  const int aex = c_lat_x[I] > 0 ? c_lat_x[I] : -c_lat_x[I];
  const int aey = c_lat_y[I] > 0 ? c_lat_y[I] : -c_lat_y[I];
  const TC vx = TC(double(aex) * in.c), vy = TC(double(aey) * in.c);
  const int dew = c_mir_ew[I], dns = c_mir_ns[I], dboth = c_mir_ns[dew];
  // this lane's column in the planes of the direction staged at a row that
  // is not (0) or is (1) mirrored north/south
  const int d0 = C.mx ? dew : I, d1 = C.mx ? dboth : dns;
  const gptr<const TS> pf0 = in.src + (long long)d0 * in.P + C.X, ph0 = in.src + (long long)(kQ + d0) * in.P + C.X;
  const gptr<const TS> pf1 = in.src + (long long)d1 * in.P + C.X, ph1 = in.src + (long long)(kQ + d1) * in.P + C.X;
  const bool nsmir = (in.ms | in.mn) != 0;
  const gptr<const uint8_t> ct = in.ct + C.X;
  const gptr<const double> al = in.alpha + C.X;
  // this lane's own cell in the planes it writes (MODE 0, 2: the dF planes
  // of direction I; MODE 1: the state planes) and, MODE 1, in the planes of
  // direction I it reads: the stored dF(f2) (FP32 storage: G, in the state
  // planes it writes)
  const gptr<TS> of = in.dst + (MODE == 1 ? (long long)I : (long long)(I - 1)) * in.P + C.X;
  const gptr<TS> oh = in.dst + (MODE == 1 ? (long long)(kQ + I) : (long long)(12 + I - 1)) * in.P + C.X;
  const gptr<const TS> df2 = kStage3G<TS> ? (gptr<const TS>)of : in.dfl + (long long)(I - 1) * in.P + C.X;
  const gptr<const TS> dh2 = kStage3G<TS> ? (gptr<const TS>)oh : in.dfl + (long long)(12 + I - 1) * in.P + C.X;

  auto map_row = [&](int yl, unsigned& my) {
    if (in.inner) {
      my = 0;
      return yl * pitch;
    }
    unsigned cy;
    return map_axis(yl, ny, in.py, in.ms, in.mn, my, cy) * pitch;
  };
  auto load = [&](int yl) {
    unsigned my;
    const int ro = map_row(prow(yl), my);
    const bool m = nsmir && my;
    SPos<TC> p;
    p.f = TC((m ? pf1 : pf0)[ro]);
    p.h = TC((m ? ph1 : ph0)[ro]);
    p.type = SOLID ? int(ct[ro]) : 0;
    p.a = SOLID ? TC(al[ro]) : TC(0);
    return p;
  };
  auto clamped_row = [&](int yl) {
    if (in.inner) return 0;
    unsigned my, cy;
    (void)map_axis(prow(yl), ny, in.py, in.ms, in.mn, my, cy);
    return (int)cy;
  };

  // x face left of this lane's column at a row: owner lane l - 1 (SX > 0) or
  // l (SX < 0); line owner - 2 sx .. owner + 2 sx across the lanes
  auto xface = [&](const SPos<TC>& r, TC& vf, TC& vh) {
    TC Lf[5], Lh[5];
    int sld[5] = {0, 0, 0, 0, 0};
    if constexpr (SX > 0) {
      Lf[3] = r.f;
      Lh[3] = r.h;
      Lf[2] = lane_shift<kDppShr1>(Lf[3]);
      Lh[2] = lane_shift<kDppShr1>(Lh[3]);
      Lf[1] = lane_shift<kDppShr1>(Lf[2]);
      Lh[1] = lane_shift<kDppShr1>(Lh[2]);
      Lf[0] = lane_shift<kDppShr1>(Lf[1]);
      Lh[0] = lane_shift<kDppShr1>(Lh[1]);
      Lf[4] = lane_shift<kDppShl1>(Lf[3]);
      Lh[4] = lane_shift<kDppShl1>(Lh[3]);
    } else {
      Lf[2] = r.f;
      Lh[2] = r.h;
      Lf[1] = lane_shift<kDppShl1>(Lf[2]);
      Lh[1] = lane_shift<kDppShl1>(Lh[2]);
      Lf[0] = lane_shift<kDppShl1>(Lf[1]);
      Lh[0] = lane_shift<kDppShl1>(Lh[1]);
      Lf[3] = lane_shift<kDppShr1>(Lf[2]);
      Lh[3] = lane_shift<kDppShr1>(Lh[2]);
      Lf[4] = lane_shift<kDppShr1>(Lf[3]);
      Lh[4] = lane_shift<kDppShr1>(Lh[3]);
    }
    if constexpr (SOLID) {
      const unsigned long long m = __ballot(r.type == 2);
      const int lane = threadIdx.x & 63;
#pragma unroll
      for (int k = 0; k < 5; k++) {
        const int pl = SX > 0 ? lane - 3 + k : lane + 2 - k;
        sld[k] = (pl >= 0 && pl < 64) ? int((m >> pl) & 1ull) : 0;
      }
      // the owner's wall closure factor, staged by the owner lane
      const TC ao = SX > 0 ? lane_shift<kDppShr1>(r.a) : r.a;
      if (sld[2]) {  // solid owner: wall closure
        strip_wall<TS>(ao, vx, Lf[2], Lh[2], rf, rh, vf, vh);
        return;
      }
    }
    strip_line<TC, SOLID, W5>(vx, Lf, Lh, sld, vf, vh);
  };

  // y face F below the window's row F (window rows F - 3 .. F + 2): owner
  // row F - 1 (SY > 0, window 2) or F (SY < 0, window 3)
  auto yface = [&](const SPos<TC> (&W)[6], TC& vf, TC& vh) {
    TC Lf[5], Lh[5];
    int sld[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const int w = SY > 0 ? k : 5 - k;
      Lf[k] = W[w].f;
      Lh[k] = W[w].h;
      sld[k] = W[w].type == 2;
    }
    if constexpr (SOLID) {
      if (sld[2]) {
        strip_wall<TS>(W[SY > 0 ? 2 : 3].a, vy, Lf[2], Lh[2], rf, rh, vf, vh);
        return;
      }
    }
    strip_line<TC, SOLID, W5>(vy, Lf, Lh, sld, vf, vh);
  };

  // Everything a row needs from memory, loaded one row ahead: the staged
  // position entering the window (row y + 3; row y itself without a y
  // component), the own cell's type as loaded (0 fluid; 1 for lanes without
  // an own cell), its row offset, (MODE 1) its stored dF(f2) (FP32
  // storage: G), and dt/dy
  // of the row (on a uniform grid the instance's own value, read with
  // stride 0, so that every row issues the same loads).
  struct Pre {
    SPos<TC> p;
    int ro, type;
    TS f, h;
    TC rdy;
  };
  auto fetch = [&](int y) {
    Pre r;
    r.p = load(HY ? y + 3 : y);
    r.ro = prow(y) * pitch;
    r.type = 1;
    r.f = r.h = TS(0);
    if (C.own) {
      r.type = ct[r.ro];
      if (MODE == 1) {
        r.f = df2[r.ro];
        r.h = dh2[r.ro];
      }
    }
    r.rdy = TC(in.rdyp[prow(y) * in.rdy_stride]);
    return r;
  };
  // a row's results, stored at the start of the next half iteration
  struct Res {
    int ro, on;
    TS vf, vh;
  };
  auto store = [&](const Res& r) {
    if (r.on) {
      of[r.ro] = r.vf;
      oh[r.ro] = r.vh;
    }
  };

  SPos<TC> W[6];  // window rows F - 3 .. F + 2 of the y face F below row F
  TC Fyf = 0, Fyh = 0;
  if constexpr (HY) {
#pragma unroll
    for (int k = 0; k < 6; k++) W[k] = load(y0 - 3 + k);
  }
  // row y from its prefetched data (registers only: no memory access)
  auto row = [&](int y, const Pre& cc) {
    // no contraction: the net flux and the state update round as the tile
    // kernels (and the oracle) round them, so both paths agree bit for bit
#pragma clang fp contract(off)
    TC Gf = 0, Gh = 0, Ff = 0, Fh = 0, Nf = 0, Nh = 0;
    SPos<TC> xr = cc.p;
    if constexpr (HY) {
#pragma unroll
      for (int k = 0; k < 5; k++) W[k] = W[k + 1];
      W[5] = cc.p;
      yface(W, Gf, Gh);  // face above row y
      xr = W[2];
    }
    if constexpr (HX) {
      xface(xr, Ff, Fh);
      Nf = lane_shift<kDppShl1>(Ff);  // face right of the column
      Nh = lane_shift<kDppShl1>(Fh);
    }
    TC oxf = 0, oxh = 0, ixf = 0, ixh = 0, oyf = 0, oyh = 0, iyf = 0, iyh = 0;
    if constexpr (HX) {
      oxf = SX > 0 ? Nf : Ff;
      oxh = SX > 0 ? Nh : Fh;
      ixf = SX > 0 ? Ff : Nf;
      ixh = SX > 0 ? Fh : Nh;
      if (SX > 0 ? C.clampL : C.clampR) {
        ixf = oxf;
        ixh = oxh;
      }
    }
    if constexpr (HY) {
      oyf = SY > 0 ? Gf : Fyf;
      oyh = SY > 0 ? Gh : Fyh;
      iyf = SY > 0 ? Fyf : Gf;
      iyh = SY > 0 ? Fyh : Gh;
      if (clamped_row(y - SY)) {
        iyf = oyf;
        iyh = oyh;
      }
      Fyf = Gf;
      Fyh = Gh;
    }
    TC dF, dH;
    if constexpr (HX && HY) {
      dF = (C.rdx * oxf + cc.rdy * oyf) - (C.rdx * ixf + cc.rdy * iyf);
      dH = (C.rdx * oxh + cc.rdy * oyh) - (C.rdx * ixh + cc.rdy * iyh);
    } else if constexpr (HX) {
      dF = C.rdx * oxf - C.rdx * ixf;
      dH = C.rdx * oxh - C.rdx * ixh;
    } else {
      dF = cc.rdy * oyf - cc.rdy * iyf;
      dH = cc.rdy * oyh - cc.rdy * iyh;
    }
    Res r;
    r.ro = cc.ro;
    r.on = cc.type == 0;
    if (MODE != 1) {
      r.vf = TS(dF);
      r.vh = TS(dH);
    } else {
      // the new state f3 + (dF2 - dF3)/2 (FP32 storage: G - dF3/2), the
      // final combination of the stage-3 kernel; the own cell's f3 is the
      // staged value of row y
      if constexpr (kStage3G<TS>) {
        r.vf = TS(TC(cc.f) - TC(0.5) * dF);
        r.vh = TS(TC(cc.h) - TC(0.5) * dH);
      } else {
        r.vf = TS(xr.f + TC(0.5) * (TC(cc.f) - dF));
        r.vh = TS(xr.h + TC(0.5) * (TC(cc.h) - dH));
      }
    }
    return r;
  };
  // Rows two at a time with two prefetch slots. Each half iteration starts
  // with one wait for every load and store in flight (on this GPU loads and
  // stores share the vector-memory counter, so any wait for a load with a
  // store pending waits for everything), which finds the previous half's
  // loads and stores complete after a whole row of arithmetic; then it
  // stores the previous row, issues the next row's loads into the other
  // slot and computes its row from registers.
  const int end = y0 + nrows;
  Pre A = fetch(y0), B;
  if constexpr (HY) {
    __builtin_amdgcn_s_waitcnt(kWaitVmem);
    yface(W, Fyf, Fyh);  // face below the first row
  }
  Res R{0, 0, TS(0), TS(0)};
  for (int y = y0;; y += 2) {
    __builtin_amdgcn_s_waitcnt(kWaitVmem);
    store(R);
    if (y + 1 < end) B = fetch(y + 1);
    R = row(y, A);
    if (y + 1 >= end) break;
    __builtin_amdgcn_s_waitcnt(kWaitVmem);
    store(R);
    if (y + 2 < end) A = fetch(y + 2);
    R = row(y + 1, B);
    if (y + 2 >= end) break;
  }
  store(R);
}

// The twelve directions of a work item, one flux method. desc: the item
// takes its rows from the last to the first in the directions with a y
// component. Items alternate between the two orders down each strip column
// (kStripDesc), so that two items one above the other reach the rows they
// share (each one's rows next to the other, read by the other as its y
// halo) at about the same time: the first read of such a row leaves it in
// L2 for the second.
template <typename TS, typename TC, int MODE, bool SOLID, bool W5>
__device__ __forceinline__ void strip_dirs(const InstDev& in, const StripInst<TS>& si, const StripLane<TC>& C, int y0,
                                           int nrows, bool desc) {
  for (int I = 1; I < kQ; I++) {
    const int sx = c_lat_x[I] > 0 ? 1 : (c_lat_x[I] < 0 ? -1 : 0);
    const int sy0 = c_lat_y[I] > 0 ? 1 : (c_lat_y[I] < 0 ? -1 : 0);
    const int fl = desc && sy0 != 0;
    const int sy = fl ? -sy0 : sy0;  // the sign in the march's row order
    // reference equilibrium of direction I for the wall faces of FP32
    // storage (a mirror image of I has the same reference: it depends on |e|)
    const TC rf = sizeof(TS) == 4 ? TC(in.ref[I]) : TC(0);
    const TC rh = sizeof(TS) == 4 ? TC(in.ref[kQ + I]) : TC(0);
    if (sy == 0) {
      if (sx > 0) strip_dir<TS, TC, MODE, SOLID, W5, 1, 0>(si, C, rf, rh, I, y0, nrows, 0);
      else strip_dir<TS, TC, MODE, SOLID, W5, -1, 0>(si, C, rf, rh, I, y0, nrows, 0);
    } else if (sx == 0) {
      if (sy > 0) strip_dir<TS, TC, MODE, SOLID, W5, 0, 1>(si, C, rf, rh, I, y0, nrows, fl);
      else strip_dir<TS, TC, MODE, SOLID, W5, 0, -1>(si, C, rf, rh, I, y0, nrows, fl);
    } else if (sy > 0) {
      if (sx > 0) strip_dir<TS, TC, MODE, SOLID, W5, 1, 1>(si, C, rf, rh, I, y0, nrows, fl);
      else strip_dir<TS, TC, MODE, SOLID, W5, -1, 1>(si, C, rf, rh, I, y0, nrows, fl);
    } else {
      if (sx > 0) strip_dir<TS, TC, MODE, SOLID, W5, 1, -1>(si, C, rf, rh, I, y0, nrows, fl);
      else strip_dir<TS, TC, MODE, SOLID, W5, -1, -1>(si, C, rf, rh, I, y0, nrows, fl);
    }
  }
}

// All twelve directions of one work item.
template <typename TS, typename TC, int MODE, bool SOLID>
__device__ __forceinline__ void strip_item(const InstDev& in, const TileRef& sr, int par) {
  const int lane = threadIdx.x & 63;
  const int x0 = sr.tx * kStripW, y0 = sr.ty, nrows = sr.flags >> 8;
  const bool desc = (sr.flags & kStripDesc) != 0;
  StripInst<TS> si;
  si.src = (gptr<const TS>)(MODE == 0 ? in.f2 : (MODE == 1 ? in.f3 : in.state[in.parity0 ^ par]));
  si.dst = (gptr<TS>)(MODE == 1 ? in.state[1 ^ in.parity0 ^ par] : in.dflux);
  si.dfl = (gptr<const TS>)in.dflux;
  si.ct = (gptr<const uint8_t>)in.ctype;
  si.alpha = (gptr<const double>)in.alpha;
  si.rdyp = (gptr<const double>)(in.rdys ? in.rdys : &in.rdy);
  si.rdy_stride = in.rdys ? 1 : 0;
  si.P = in.plane;
  si.pitch = in.pitch;
  si.ny = in.ny;
  si.py = in.flags & kPeriodicY;
  si.ms = in.flags & kMirrorS;
  si.mn = in.flags & kMirrorN;
  si.inner = y0 - kH >= 0 && y0 + nrows + kH <= in.ny;
  si.c = in.c;
  StripLane<TC> C;
  C.xl = x0 - kH + lane;
  const int px = in.flags & kPeriodicX, mw = in.flags & kMirrorW, me = in.flags & kMirrorE;
  unsigned mx, cx, mt, cl, cr;
  C.X = map_axis(C.xl, in.nx, px, mw, me, mx, cx);
  C.mx = (int)mx;
  (void)map_axis(C.xl - 1, in.nx, px, mw, me, mt, cl);
  (void)map_axis(C.xl + 1, in.nx, px, mw, me, mt, cr);
  C.clampL = (int)cl;
  C.clampR = (int)cr;
  C.own = lane >= kH && lane < kH + kStripW && C.xl < in.nx;
  C.rdx = C.own ? TC(in.rdxs ? in.rdxs[C.xl] : in.rdx) : TC(0);
  if (in.flux_method == 1) strip_dirs<TS, TC, MODE, SOLID, true>(in, si, C, y0, nrows, desc);
  else strip_dirs<TS, TC, MODE, SOLID, false>(in, si, C, y0, nrows, desc);
}

// Persistent grid of 4-wave blocks; wave w of the (XCD-renumbered) grid
// takes items w, w + 4 G, ... of the list. At most 168 registers (3 waves
// per SIMD): capped at 128 for a fourth wave the FP64 kernels spill (up to
// 73 slots) and run slower (4096^2 Riemann stage-2 and stage-3 flux 2.19
// and 2.82 ms at 4 waves, 1.77 and 2.18 ms at 3).
template <typename TS, typename TC, int MODE, bool SOLID>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_strip(
    const InstDev* __restrict__ insts, const TileRef* __restrict__ strips, int nstrips, int par, const int* stepp) {
  const int step = *stepp;
  const int wave = __builtin_amdgcn_readfirstlane(first_tile() * 4 + (int)(threadIdx.x >> 6));
  const int nw = gridDim.x * 4;
  for (int it = wave; it < nstrips; it += nw) {
    const TileRef sr = strips[it];
    const InstDev& in = insts[sr.inst];
    if (frozen_by_failure(in, step)) continue;
    if (!eq_kernel_for<TS, TC>(in)) continue;  // FP32 storage: the other compute type's instance
    strip_item<TS, TC, MODE, SOLID>(in, sr, par);
  }
}
